// store.hip -- on-disk / wire format of a built H-matrix (no counterpart in the
// reference, whose only persistence is text dumps for debugging, tree.cu:173-183,
// helper.cu:203-255).
//
// A file holds what the engine cannot recompute cheaply: the ACA factors (FP64 or
// the mixed-precision BF16 panels) and their effective ranks, plus everything the
// deterministic, bit-exact setup needs to rebuild the identical block list and
// layout: the parameters, the kernel and the points in the caller's order.
// Loading runs setup (Morton, sort, tree, schedule: a fraction of the ACA's cost),
// checks that the layout sizes equal the stored ones, and installs the factors;
// products are then bitwise those of the saved H-matrix.
//
// Layout (little-endian), all sizes in bytes:
//   StoreHeader (fixed size, below)
//   points of set 0: dim arrays of n0 doubles (caller's order)
//   points of set 1 (absent when the header flags the same points)
//   effective ranks: naca int32
//   U, then V: u_elems and v_elems elements of elem_bytes each (8: FP64 layout,
//   2: BF16 hi + lo panels)
// The header's checksum covers the header itself (with its checksum field zero)
// and the payload (64-bit multiply-xor hash over 8-byte words, the tail
// zero-padded), so a damaged parameter or kernel field is refused like a damaged
// factor.
#include <cstdio>
#include <cstring>
#include <vector>

#include "hmg_internal.h"
#include "kernel_system_assembler.h"

namespace hmg {
namespace {

constexpr char kMagic[8] = {'H', 'M', 'G', 'A', 'M', 'D', 'H', 'M'};
constexpr uint32_t kVersion = 2;

struct StoreHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int32_t dim, bits, n0, n1;
    int32_t max_level, c_leaf, k, precision;
    int32_t root_level_1, root_level_2, kernel_kind, same_points;
    int32_t near_symmetric, pad0;
    double eta, epsilon, regularization, length_scale;
    int64_t nblocks, ndense, naca;
    int64_t u_elems, v_elems, elem_bytes;
    uint64_t checksum;
};

struct Hash {
    uint64_t h = 0x9e3779b97f4a7c15ull;
    unsigned char tail[8];
    int ntail = 0;
    void word(uint64_t w)
    {
        h ^= w;
        h *= 0x100000001b3ull;
        h ^= h >> 29;
    }
    void add(const void* p, size_t n)
    {
        const unsigned char* b = (const unsigned char*)p;
        while (n > 0 && ntail > 0) {
            tail[ntail++] = *b++;
            n--;
            if (ntail == 8) {
                uint64_t w;
                memcpy(&w, tail, 8);
                word(w);
                ntail = 0;
            }
        }
        for (; n >= 8; n -= 8, b += 8) {
            uint64_t w;
            memcpy(&w, b, 8);
            word(w);
        }
        while (n > 0) {
            tail[ntail++] = *b++;
            n--;
        }
    }
    uint64_t finish()
    {
        if (ntail > 0) {
            memset(tail + ntail, 0, 8 - ntail);
            uint64_t w;
            memcpy(&w, tail, 8);
            word(w);
            ntail = 0;
        }
        return h;
    }
};

constexpr size_t kChunk = (size_t)64 << 20;

struct File {
    FILE* f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

// device -> file in chunks, hashing what is written
hmg_status write_device(FILE* f, const void* dev, size_t bytes, hipStream_t s, std::vector<char>& buf,
                        Hash& hash)
{
    for (size_t off = 0; off < bytes; off += kChunk) {
        const size_t n = std::min(kChunk, bytes - off);
        buf.resize(n);
        HMG_HIP(hipMemcpyAsync(buf.data(), (const char*)dev + off, n, hipMemcpyDeviceToHost, s));
        HMG_HIP(hipStreamSynchronize(s));
        hash.add(buf.data(), n);
        if (fwrite(buf.data(), 1, n, f) != n) {
            set_error("write failed");
            return HMG_ERR_INVALID;
        }
    }
    return HMG_OK;
}

hmg_status read_device(FILE* f, void* dev, size_t bytes, hipStream_t s, std::vector<char>& buf, Hash& hash)
{
    for (size_t off = 0; off < bytes; off += kChunk) {
        const size_t n = std::min(kChunk, bytes - off);
        buf.resize(n);
        if (fread(buf.data(), 1, n, f) != n) {
            set_error("file truncated");
            return HMG_ERR_INVALID;
        }
        hash.add(buf.data(), n);
        HMG_HIP(hipMemcpyAsync((char*)dev + off, buf.data(), n, hipMemcpyHostToDevice, s));
        HMG_HIP(hipStreamSynchronize(s));
    }
    return HMG_OK;
}

hmg_status write_host(FILE* f, const void* p, size_t bytes, Hash& hash)
{
    hash.add(p, bytes);
    if (bytes && fwrite(p, 1, bytes, f) != bytes) {
        set_error("write failed");
        return HMG_ERR_INVALID;
    }
    return HMG_OK;
}

hmg_status read_host(FILE* f, void* p, size_t bytes, Hash& hash)
{
    if (bytes && fread(p, 1, bytes, f) != bytes) {
        set_error("file truncated");
        return HMG_ERR_INVALID;
    }
    hash.add(p, bytes);
    return HMG_OK;
}

hmg_status read_header(FILE* f, StoreHeader* h)
{
    if (fread(h, 1, sizeof(*h), f) != sizeof(*h)) {
        set_error("not an H-matrix file (short header)");
        return HMG_ERR_INVALID;
    }
    if (memcmp(h->magic, kMagic, 8) != 0) {
        set_error("not an H-matrix file (magic)");
        return HMG_ERR_INVALID;
    }
    if (h->version != kVersion || h->header_bytes != sizeof(StoreHeader)) {
        set_error("H-matrix file version " + std::to_string(h->version) + " is not supported");
        return HMG_ERR_INVALID;
    }
    if (h->dim < 1 || h->dim > MAX_POINT_DIMENSION || h->n0 < 1 || h->n1 < 1 || h->naca < 0 ||
        (h->precision != HMG_PREC_FP64 && h->precision != HMG_PREC_MIXED) ||
        h->elem_bytes != (h->precision == HMG_PREC_MIXED ? 2 : 8) || h->u_elems < 0 || h->v_elems < 0) {
        set_error("H-matrix file header is inconsistent");
        return HMG_ERR_INVALID;
    }
    return HMG_OK;
}

// the points of set `set` in the caller's order: sorted coordinates scattered
// back through the sort permutation
hmg_status original_points(const h_matrix_data* d, int set, hipStream_t s, std::vector<double>& out)
{
    const int n = d->point_count[set], dim = d->dim;
    std::vector<uint64_t> order(n);
    HMG_HIP(hipMemcpyAsync(order.data(), d->order[set], sizeof(uint64_t) * n, hipMemcpyDeviceToHost, s));
    std::vector<double> sorted(n);
    out.assign((size_t)dim * n, 0.0);
    for (int k = 0; k < dim; k++) {
        HMG_HIP(hipMemcpyAsync(sorted.data(), d->coords_d[set][k], sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HMG_HIP(hipStreamSynchronize(s));
        for (int p = 0; p < n; p++) out[(size_t)k * n + order[p]] = sorted[p];
    }
    return HMG_OK;
}

}  // namespace
}  // namespace hmg

using namespace hmg;

extern "C" {

hmg_status hmg_save(const struct h_matrix_data* data, const char* path)
{
    Engine* e = engine_of(data);
    if (!e || !e->built) {
        set_error("H-matrix not set up");
        return HMG_ERR_NOT_READY;
    }
    if (!path) {
        set_error("null path");
        return HMG_ERR_INVALID;
    }
    if (data->U == nullptr || e->aca_k < 0) {
        set_error("no precomputed ACA factors to save (call precompute_aca first)");
        return HMG_ERR_NOT_READY;
    }
    if (!data->assem) {
        set_error("h_matrix_data.assem is not set");
        return HMG_ERR_NOT_READY;
    }
    system_assembler a{};
    hipStream_t s = e->stream;
    HMG_HIP(hipMemcpyAsync(&a, data->assem, sizeof(a), hipMemcpyDeviceToHost, s));
    HMG_HIP(hipStreamSynchronize(s));
    const bool mixed = e->fact_prec == HMG_PREC_MIXED;
    const int na = e->naca;

    std::vector<double> p0, p1;
    HMG_TRY(original_points(data, 0, s, p0));
    HMG_TRY(original_points(data, 1, s, p1));
    // bitwise, as setup decides it (value equality would take -0.0 for +0.0)
    const bool same = e->n[0] == e->n[1] && memcmp(p0.data(), p1.data(), sizeof(double) * p0.size()) == 0;

    StoreHeader h{};
    memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.header_bytes = sizeof(StoreHeader);
    h.dim = data->dim;
    h.bits = data->bits;
    h.n0 = e->n[0];
    h.n1 = e->n[1];
    h.max_level = data->max_level;
    h.c_leaf = data->c_leaf;
    h.k = e->aca_k;
    h.precision = e->fact_prec;
    h.root_level_1 = (int)data->root_level_set_1;
    h.root_level_2 = (int)data->root_level_set_2;
    h.kernel_kind = a.kind;
    h.same_points = same ? 1 : 0;
    h.near_symmetric = e->near_sym_enabled ? 1 : 0;
    h.eta = data->eta;
    h.epsilon = data->epsilon;
    h.regularization = a.regularization;
    h.length_scale = a.length_scale;
    h.nblocks = e->nblocks;
    h.ndense = e->ndense;
    h.naca = na;
    h.u_elems = mixed ? (int64_t)e->Up.n : (int64_t)e->U.n;
    h.v_elems = mixed ? (int64_t)e->Vp.n : (int64_t)e->V.n;
    h.elem_bytes = mixed ? 2 : 8;

    File f;
    f.f = fopen(path, "wb");
    if (!f.f) {
        set_error(std::string("cannot open ") + path + " for writing");
        return HMG_ERR_INVALID;
    }
    // header with a zero checksum first, the real one written at the end
    if (fwrite(&h, 1, sizeof(h), f.f) != sizeof(h)) {
        set_error("write failed");
        return HMG_ERR_INVALID;
    }
    Hash hash;
    hash.add(&h, sizeof(h));  // checksum field still zero
    HMG_TRY(write_host(f.f, p0.data(), sizeof(double) * p0.size(), hash));
    if (!same) HMG_TRY(write_host(f.f, p1.data(), sizeof(double) * p1.size(), hash));
    std::vector<int> keff(std::max(na, 1), 0);
    if (na > 0) {
        HMG_HIP(hipMemcpyAsync(keff.data(), e->keff.p, sizeof(int) * na, hipMemcpyDeviceToHost, s));
        HMG_HIP(hipStreamSynchronize(s));
    }
    HMG_TRY(write_host(f.f, keff.data(), sizeof(int) * na, hash));
    std::vector<char> buf;
    const void* U = mixed ? (const void*)e->Up.p : (const void*)e->U.p;
    const void* V = mixed ? (const void*)e->Vp.p : (const void*)e->V.p;
    HMG_TRY(write_device(f.f, U, (size_t)h.u_elems * h.elem_bytes, s, buf, hash));
    HMG_TRY(write_device(f.f, V, (size_t)h.v_elems * h.elem_bytes, s, buf, hash));
    h.checksum = hash.finish();
    if (fseek(f.f, 0, SEEK_SET) != 0 || fwrite(&h, 1, sizeof(h), f.f) != sizeof(h) || fflush(f.f) != 0) {
        set_error("write failed");
        return HMG_ERR_INVALID;
    }
    return HMG_OK;
}

hmg_status hmg_load(struct h_matrix_data* data, const char* path)
{
    if (!data || !path) {
        set_error("null argument");
        return HMG_ERR_INVALID;
    }
    File f;
    f.f = fopen(path, "rb");
    if (!f.f) {
        set_error(std::string("cannot open ") + path);
        return HMG_ERR_INVALID;
    }
    StoreHeader h;
    HMG_TRY(read_header(f.f, &h));
    // the host part of the payload first: nothing touches the GPU before it checks out
    Hash hash;
    {
        StoreHeader h0 = h;
        h0.checksum = 0;
        hash.add(&h0, sizeof(h0));
    }
    std::vector<double> p0((size_t)h.dim * h.n0), p1;
    HMG_TRY(read_host(f.f, p0.data(), sizeof(double) * p0.size(), hash));
    if (!h.same_points) {
        p1.resize((size_t)h.dim * h.n1);
        HMG_TRY(read_host(f.f, p1.data(), sizeof(double) * p1.size(), hash));
    }
    std::vector<int> keff(std::max<int64_t>(h.naca, 1), 0);
    HMG_TRY(read_host(f.f, keff.data(), sizeof(int) * h.naca, hash));

    int pc[2] = {h.n0, h.n1};
    HMG_TRY(hmg_init_h_matrix_data(data, pc, h.dim, h.bits));
    Engine* e = engine_of(data);
    hipStream_t s = e->stream;
    auto fail = [&](hmg_status st) {
        (void)hmg_destroy_h_matrix_data(data);
        return st;
    };
    for (int i = 0; i < 2; i++) {
        const std::vector<double>& p = (i == 1 && !h.same_points) ? p1 : p0;
        for (int k = 0; k < h.dim; k++)
            if (hipMemcpyAsync(data->coords_d[i][k], p.data() + (size_t)k * pc[i], sizeof(double) * pc[i],
                               hipMemcpyHostToDevice, s) != hipSuccess) {
                set_error("coordinate upload failed");
                return fail(HMG_ERR_HIP);
            }
    }
    if (hipStreamSynchronize(s) != hipSuccess) return fail(HMG_ERR_HIP);
    data->eta = h.eta;
    data->epsilon = h.epsilon;
    data->max_level = h.max_level;
    data->c_leaf = h.c_leaf;
    data->k = h.k;
    data->root_level_set_1 = (char)h.root_level_1;
    data->root_level_set_2 = (char)h.root_level_2;
    data->max_batched_dense_size = 1 << 20;
    data->max_batched_aca_size = 1 << 30;
    e->near_sym_enabled = h.near_symmetric != 0;
    hmg_status st = hmg_setup_h_matrix(data);
    if (st != HMG_OK) return fail(st);
    if (e->nblocks != h.nblocks || e->ndense != h.ndense || e->naca != h.naca) {
        set_error("the rebuilt block list differs from the stored one");
        return fail(HMG_ERR_INVALID);
    }
    system_assembler* kern = hmg_kernel_new(h.kernel_kind, h.regularization, h.length_scale);
    if (!kern) return fail(HMG_ERR_INVALID);
    e->own_kernel = kern;
    data->assem = kern;
    e->precision = h.precision;
    AllocScope scope(s);
    const bool mixed = h.precision == HMG_PREC_MIXED;
    st = aca_layout(*e, h.k, h.precision);
    if (st != HMG_OK) return fail(st);
    const int64_t nu = mixed ? e->up_size : e->h_uoff[e->naca] + 256;
    const int64_t nv = mixed ? e->vp_size : e->h_voff[e->naca] + 256;
    if (nu != h.u_elems || nv != h.v_elems) {
        set_error("the rebuilt factor layout differs from the stored one");
        return fail(HMG_ERR_INVALID);
    }
    std::vector<char> buf;
    void *U = nullptr, *V = nullptr;
    if (mixed) {
        if ((st = e->Up.alloc((size_t)nu)) != HMG_OK || (st = e->Vp.alloc((size_t)nv)) != HMG_OK) return fail(st);
        U = e->Up.p;
        V = e->Vp.p;
    } else {
        if ((st = e->U.alloc((size_t)nu)) != HMG_OK || (st = e->V.alloc((size_t)nv)) != HMG_OK) return fail(st);
        U = e->U.p;
        V = e->V.p;
    }
    if ((st = read_device(f.f, U, (size_t)nu * h.elem_bytes, s, buf, hash)) != HMG_OK ||
        (st = read_device(f.f, V, (size_t)nv * h.elem_bytes, s, buf, hash)) != HMG_OK)
        return fail(st);
    if (hash.finish() != h.checksum) {
        set_error("H-matrix file checksum mismatch (corrupted payload)");
        return fail(HMG_ERR_INVALID);
    }
    if (e->naca > 0 && hipMemcpyAsync(e->keff.p, keff.data(), sizeof(int) * e->naca, hipMemcpyHostToDevice, s) !=
                           hipSuccess)
        return fail(HMG_ERR_HIP);
    if (hipStreamSynchronize(s) != hipSuccess) return fail(HMG_ERR_HIP);
    data->U = (double*)U;
    data->V = (double*)V;
    e->fact_prec = h.precision;
    e->epoch++;
    return HMG_OK;
}

}  // extern "C"
