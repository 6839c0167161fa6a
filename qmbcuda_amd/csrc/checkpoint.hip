// On-disk ELL format: save a device-built matrix and load it back for the same
// basis (restartable runs, reuse by the CPU baseline). The reference only has a
// commented-out raw dump of its value and index arrays as float
// (Q/exact_diagonalization/HeisenbergHamAbelianSymms_CUDA.cu:149-162 with
// WriteArrayToFile, Q/utils/print_funcs.cu:73-90); this format keeps the exact
// device layout (SELL-64, FP64 or packed values) behind a versioned header that
// ties the file to its basis.
//
// Layout (little-endian): EllFileHeader, then chunk_ptr (int64, n_chunks + 1),
// row_len (uint16, n_rows), col (uint32, stored), then either val (double,
// stored x 1 or 2) or, packed, diag (double, n_rows x 1 or 2), the value table
// (double, n_values) and, complex, the imaginary-part table (double,
// n_values_im).
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "checkpoint.hpp"

namespace qmb {

namespace {

struct EllFileHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int32_t value_type, packed, col_bits, n_values, max_terms, max_row_len;
    int64_t n_rows, n_chunks, nnz, stored;
    int32_t n_sites, n_up, irrep, layout;
    uint64_t group_hash;
    int32_t re_bits, n_values_im;  // packed complex values
};

constexpr char kEllMagic[8] = {'Q', 'M', 'B', 'E', 'L', 'L', '0', '1'};

}  // namespace

File::File(const char *path, const char *mode) : f(std::fopen(path, mode)), name(path) {
    QMB_REQUIRE(f != nullptr, std::string("cannot open ") + path);
}
File::~File() {
    if (f) std::fclose(f);
}
void File::write(const void *p, size_t n) {
    QMB_REQUIRE(std::fwrite(p, 1, n, f) == n, "write failed: " + name);
}
void File::read(void *p, size_t n) {
    QMB_REQUIRE(std::fread(p, 1, n, f) == n, "file too short: " + name);
}
void File::close() {
    if (f) {
        const int rc = std::fclose(f);
        f = nullptr;
        QMB_REQUIRE(rc == 0, "close failed: " + name);
    }
}

void write_device(File &out, const void *dev, size_t bytes) {
    if (bytes == 0) return;
    const size_t chunk = std::min<size_t>(bytes, kStageBytes);
    PinnedBuf stage(chunk);
    for (size_t off = 0; off < bytes; off += chunk) {
        const size_t n = std::min(chunk, bytes - off);
        QMB_HIP(hipMemcpy(stage.p, static_cast<const char *>(dev) + off, n, hipMemcpyDeviceToHost));
        out.write(stage.p, n);
    }
}

void read_device(File &in, void *dev, size_t bytes) {
    if (bytes == 0) return;
    const size_t chunk = std::min<size_t>(bytes, kStageBytes);
    PinnedBuf stage(chunk);
    for (size_t off = 0; off < bytes; off += chunk) {
        const size_t n = std::min(chunk, bytes - off);
        in.read(stage.p, n);
        QMB_HIP(hipMemcpy(static_cast<char *>(dev) + off, stage.p, n, hipMemcpyHostToDevice));
    }
}

// Bounds of a loaded matrix, checked on the device before any SpMV may read it:
// chunk offsets ascending in whole slots, every row within its chunk's width and
// max_terms, every stored column below n_rows (and, packed, every value index
// inside the table). err = 1 on the first violation.
__global__ __launch_bounds__(256) void ell_validate(int64_t n_rows, int64_t n_chunks, const int64_t *chunk_ptr,
                                                     const uint16_t *row_len, const uint32_t *col, int64_t stored,
                                                     int packed, int cb, int n_values, int cplx, int rb,
                                                     int n_values_im, int W, int *err) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_chunks * 64;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = r / 64;
        const int lane = (int)(r % 64);
        const int64_t base = chunk_ptr[c], end = chunk_ptr[c + 1];
        if (base < 0 || end < base || end > stored || (end - base) % 64) {
            *err = 1;
            continue;
        }
        const int L = (int)((end - base) / 64);
        const int len = r < n_rows ? (int)row_len[r] : 0;
        if (len > L || len > W) {
            *err = 1;
            continue;
        }
        const uint32_t mask = packed ? ((1u << cb) - 1u) : 0xFFFFFFFFu;
        for (int k = packed ? 1 : 0; k < len; ++k) {
            const uint32_t w = col[base + (int64_t)k * 64 + lane];
            if ((int64_t)(w & mask) >= n_rows) *err = 1;
            if (packed && !cplx && (int)(w >> cb) >= n_values) *err = 1;
            if (packed && cplx &&
                ((int)((w >> cb) & ((1u << rb) - 1u)) >= n_values ||
                 (int)((uint64_t)w >> (cb + rb)) >= n_values_im))
                *err = 1;
        }
    }
}

}  // namespace qmb

using namespace qmb;

extern "C" {

qmb_status qmb_ell_save(const qmb_ell *e, const char *path) {
    return guarded([&] {
        QMB_REQUIRE(e && path, "qmb_ell_save: NULL argument");
        QMB_HIP(hipSetDevice(e->ctx->device));
        QMB_HIP(hipStreamSynchronize(e->ctx->stream));
        const qmb_basis *b = e->basis;
        EllFileHeader h{};
        std::memcpy(h.magic, kEllMagic, 8);
        h.version = 2;
        h.header_bytes = sizeof(EllFileHeader);
        h.value_type = e->value_type;
        h.packed = e->packed;
        h.col_bits = e->col_bits;
        h.n_values = e->n_values;
        h.max_terms = e->max_terms;
        h.max_row_len = e->max_row_len;
        h.n_rows = e->n_rows;
        h.n_chunks = e->n_chunks;
        h.nnz = e->nnz;
        h.stored = e->stored;
        h.n_sites = b->n_sites;
        h.n_up = b->n_up;
        h.irrep = b->irrep;
        h.layout = b->layout;
        h.group_hash = b->group->hash;
        h.re_bits = e->re_bits;
        h.n_values_im = e->n_values_im;
        File out(path, "wb");
        out.write(&h, sizeof h);
        const int vs = e->value_type == QMB_COMPLEX ? 2 : 1;
        write_device(out, e->chunk_ptr.p, (size_t)(e->n_chunks + 1) * sizeof(int64_t));
        write_device(out, e->row_len.p, (size_t)e->n_rows * sizeof(uint16_t));
        write_device(out, e->col.p, (size_t)e->stored * sizeof(int32_t));
        if (e->packed) {
            write_device(out, e->diag.p, (size_t)e->n_rows * vs * sizeof(double));
            write_device(out, e->vtab.p, (size_t)e->n_values * sizeof(double));
            if (vs == 2) write_device(out, e->vtab.as<double>() + kMaxValues, (size_t)e->n_values_im * sizeof(double));
        } else {
            write_device(out, e->val.p, (size_t)e->stored * vs * sizeof(double));
        }
        out.close();
    });
}

qmb_status qmb_ell_load(qmb_ctx *ctx, const qmb_basis *b, const char *path, qmb_ell **out) {
    return guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(out, "qmb_ell_load: out is NULL");
        *out = nullptr;
        QMB_REQUIRE(ctx && b && path, "qmb_ell_load: NULL argument");
        QMB_REQUIRE(b->ctx == ctx, "qmb_ell_load: basis belongs to another context");
        QMB_HIP(hipSetDevice(ctx->device));
        File in(path, "rb");
        EllFileHeader h{};
        in.read(&h, sizeof h);
        QMB_REQUIRE(std::memcmp(h.magic, kEllMagic, 8) == 0, "qmb_ell_load: not an ELL file");
        QMB_REQUIRE(h.version == 2 && h.header_bytes == sizeof(EllFileHeader), "qmb_ell_load: unknown version");
        QMB_REQUIRE(h.n_rows == b->n_rows && h.n_sites == b->n_sites && h.n_up == b->n_up && h.irrep == b->irrep &&
                        h.layout == b->layout && h.group_hash == b->group->hash,
                    "qmb_ell_load: the file was built for another basis");
        QMB_REQUIRE((h.value_type == QMB_REAL || h.value_type == QMB_COMPLEX) && h.n_chunks == (h.n_rows + 63) / 64 &&
                        h.stored >= 0 && h.n_values >= 0 && h.n_values <= kMaxValues && h.n_values_im >= 0 &&
                        h.n_values_im <= kMaxValues && h.re_bits >= 0 && h.re_bits <= 31,
                    "qmb_ell_load: corrupt header");
        auto e = std::make_unique<qmb_ell>();
        e->ctx = ctx;
        e->basis = b;
        e->value_type = h.value_type;
        e->max_terms = h.max_terms;
        e->n_rows = h.n_rows;
        e->n_chunks = h.n_chunks;
        e->nnz = h.nnz;
        e->stored = h.stored;
        e->max_row_len = h.max_row_len;
        e->packed = h.packed;
        e->col_bits = h.col_bits;
        e->n_values = h.n_values;
        e->re_bits = h.re_bits;
        e->n_values_im = h.n_values_im;
        const int vs = h.value_type == QMB_COMPLEX ? 2 : 1;
        e->chunk_ptr.alloc((size_t)(h.n_chunks + 1) * sizeof(int64_t));
        e->row_len.alloc((size_t)std::max<int64_t>(h.n_rows, 1) * sizeof(uint16_t));
        e->col.alloc((size_t)std::max<int64_t>(h.stored, 1) * sizeof(int32_t));
        read_device(in, e->chunk_ptr.p, (size_t)(h.n_chunks + 1) * sizeof(int64_t));
        read_device(in, e->row_len.p, (size_t)h.n_rows * sizeof(uint16_t));
        read_device(in, e->col.p, (size_t)h.stored * sizeof(int32_t));
        if (h.packed) {
            e->diag.alloc((size_t)std::max<int64_t>(h.n_rows, 1) * vs * sizeof(double));
            e->vtab.alloc((size_t)kMaxValues * vs * sizeof(double));
            read_device(in, e->diag.p, (size_t)h.n_rows * vs * sizeof(double));
            read_device(in, e->vtab.p, (size_t)h.n_values * sizeof(double));
            if (vs == 2) read_device(in, e->vtab.as<double>() + kMaxValues, (size_t)h.n_values_im * sizeof(double));
        } else {
            e->val.alloc((size_t)std::max<int64_t>(h.stored, 1) * vs * sizeof(double));
            read_device(in, e->val.p, (size_t)h.stored * vs * sizeof(double));
        }
        // the offsets must start at 0 and end at the stored count, and every entry
        // an SpMV reads must lie inside the arrays
        int64_t first = -1, last = 0;
        QMB_HIP(hipMemcpy(&first, e->chunk_ptr.p, sizeof first, hipMemcpyDeviceToHost));
        QMB_HIP(hipMemcpy(&last, e->chunk_ptr.as<int64_t>() + h.n_chunks, sizeof last, hipMemcpyDeviceToHost));
        QMB_REQUIRE(first == 0 && last == h.stored, "qmb_ell_load: corrupt chunk offsets");
        const bool cplx = h.value_type == QMB_COMPLEX;
        QMB_REQUIRE(!h.packed || (h.col_bits >= 1 && h.col_bits <= 31 && ((int64_t)1 << h.col_bits) >= h.n_rows &&
                                  h.n_values >= 1 &&
                                  (cplx ? (h.n_values_im >= 1 && h.re_bits == index_bits(h.n_values) &&
                                           h.col_bits + h.re_bits + index_bits(h.n_values_im) <= 32)
                                        : h.n_values <= ((int64_t)1 << (32 - h.col_bits)))),
                    "qmb_ell_load: corrupt packing parameters");
        DevBuf err(sizeof(int));
        QMB_HIP(hipMemsetAsync(err.p, 0, sizeof(int), ctx->stream));
        if (h.n_rows > 0) {
            const unsigned blocks = (unsigned)std::min<int64_t>(4096, (h.n_chunks * 64 + 255) / 256);
            hipLaunchKernelGGL(ell_validate, dim3(blocks), dim3(256), 0, ctx->stream, h.n_rows, h.n_chunks,
                               e->chunk_ptr.as<int64_t>(), e->row_len.as<uint16_t>(), e->col.as<uint32_t>(),
                               h.stored, h.packed, h.col_bits, h.n_values, cplx ? 1 : 0, h.re_bits, h.n_values_im,
                               h.max_terms, err.as<int>());
            QMB_HIP(hipGetLastError());
        }
        int herr = 0;
        QMB_HIP(hipMemcpyAsync(&herr, err.p, sizeof herr, hipMemcpyDeviceToHost, ctx->stream));
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        QMB_REQUIRE(herr == 0, "qmb_ell_load: the matrix indexes outside its arrays (corrupt file)");
        *out = e.release();
    });
}

}  // extern "C"
