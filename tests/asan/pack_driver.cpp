// Host-side checks of the C ABI's validation and table packing under AddressSanitizer: every
// host table is allocated at exactly the size the library asks for, so a write past it, a read of
// a descriptor beyond `count` or a use of freed memory stops the run.  No GPU is touched: the
// descriptors carry made-up device addresses, which the packers copy and never dereference.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/tfdtd.h"

namespace {

int checks = 0, failures = 0;

void expect(bool ok, const char* what) {
    ++checks;
    if (!ok) {
        ++failures;
        char message[256] = {0};
        tfdtd_last_error(message, sizeof message);
        std::fprintf(stderr, "FAILED: %s (last error: %s)\n", what, message);
    }
}

void* address(uintptr_t k) { return reinterpret_cast<void*>(0x10000000u + k * 0x1000u); }

// A family with `layers`-row slabs on both faces of every axis, psi arrays laid out one after
// another and one coefficient block per face.
tfdtd_march_family family(const int n[3], int layers, bool split, uintptr_t base) {
    tfdtd_march_family f;
    std::memset(&f, 0, sizeof f);
    f.psi_in = address(base);
    f.psi_out = split ? address(base + 1) : address(base);
    f.coef = address(base + 2);
    int64_t psi = 0, coef = 0;
    for (int ax = 0; ax < 3; ++ax) {
        const int64_t plane = (int64_t)n[0] * n[1] * n[2] / n[ax];
        for (int side = 0; side < 2; ++side) {
            tfdtd_march_face& face = f.face[ax][side];
            face.lo = side ? n[ax] - layers : 1;
            face.hi = side ? n[ax] : layers;
            const int64_t w = face.hi - face.lo;
            face.psi[0] = psi;
            face.psi[1] = psi + w * plane;
            psi += 2 * w * plane;
            face.coef = coef;
            coef += 3 * w;
        }
    }
    return f;
}

tfdtd_march march(int nx, int ny, int nz, int32_t dtype, uintptr_t base, int layers = 3) {
    tfdtd_march d;
    std::memset(&d, 0, sizeof d);
    d.dtype = dtype;
    d.nx = nx; d.ny = ny; d.nz = nz;
    const int n[3] = {nx, ny, nz};
    d.e_in = address(base); d.h_in = address(base + 1);
    d.e_out = address(base + 2); d.h_out = address(base + 3);
    d.inverse_permittivity = address(base + 4);
    d.eps_stride = 1;
    d.inverse_permeability = address(base + 5);
    d.courant = 0.5;
    d.e = family(n, layers, true, base + 10);
    d.h = family(n, layers, false, base + 20);     // the one-step pass updates the H memory in place
    d.nsrc_e = 1;
    d.srcs_e[0].comp = 2;
    for (int ax = 0; ax < 3; ++ax) { d.srcs_e[0].lo[ax] = n[ax] / 2; d.srcs_e[0].hi[ax] = n[ax] / 2 + 1; }
    d.srcs_e[0].values = address(base + 30);
    d.source_samples = 100;
    d.counter = static_cast<const int64_t*>(address(base + 31));
    d.tile_y = 8;
    d.chunk_x = 6;
    return d;
}

tfdtd_half_step half_step(int nx, int ny, int nz, uintptr_t base) {
    tfdtd_half_step d;
    std::memset(&d, 0, sizeof d);
    d.dtype = TFDTD_F32;
    d.family = TFDTD_FAMILY_E;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.src = address(base); d.dst = address(base + 1); d.inverse = address(base + 2);
    d.inverse_stride = 1;
    d.courant = 0.5;
    d.nseg = 2;
    for (int k = 0; k < 2; ++k) {
        d.seg[k].axis = 0; d.seg[k].comp = 1 + k; d.seg[k].lo = 1; d.seg[k].hi = 4;
        d.seg[k].psi = address(base + 3 + k); d.seg[k].b = address(base + 5);
        d.seg[k].c = address(base + 6); d.seg[k].inv_k = address(base + 7);
    }
    d.counter = static_cast<const int64_t*>(address(base + 8));
    return d;
}

template <typename Pack>
int pack_exact(size_t bytes, Pack&& pack) {
    void* table = std::malloc(bytes ? bytes : 1);
    const int status = pack(table, bytes);
    std::free(table);
    return status;
}

}  // namespace

int main() {
    // One-pass tables of grids of three shapes, both precisions.
    for (int32_t dtype : {TFDTD_F32, TFDTD_F64}) {
        std::vector<tfdtd_march> grids = {march(32, 24, 20, dtype, 100), march(17, 40, 64, dtype, 200),
                                          march(64, 9, 70, dtype, 300)};
        const int count = (int)grids.size();
        const size_t bytes = tfdtd_batch_march_bytes(count, dtype);
        std::vector<int32_t> prefix(3 * (count + 1));
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march_pack(grids.data(), count, t, b, prefix.data()); }) == TFDTD_OK,
               "batch march pack of three grids");
        expect(prefix[count] + prefix[2 * count + 1] + prefix[3 * count + 2] > 0, "the prefixes count workgroups");
        expect(pack_exact(bytes - 1, [&](void* t, size_t b) { return tfdtd_batch_march_pack(grids.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "a table one byte short is refused");
        std::vector<tfdtd_march> bad = grids;
        bad[1].e.face[0][1].hi = bad[1].nx + 1;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "a slab past the axis is refused");
        bad = grids;
        bad[2].e.face[1][0].hi = bad[2].e.face[1][1].lo + 1;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "overlapping faces are refused");
        bad = grids;
        bad[0].nsrc_e = TFDTD_MAX_SOURCES + 1;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "too many sources are refused");
        bad = grids;
        bad[0].srcs_e[0].hi[2] = bad[0].nz + 1;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "a source box past the grid is refused");
        bad = grids;
        bad[0].counter = nullptr;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "a missing counter is refused");
        bad = grids;
        bad[1].snap_e = bad[1].snap_h = address(910);
        bad[1].snap_count = 1;
        for (int ax = 0; ax < 3; ++ax) { bad[1].snap_lo[0][ax] = 2; bad[1].snap_hi[0][ax] = 4; }
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "a step n+1 snapshot is refused by the one-step kernel");
    }
    // Two-step tables (float32, both families double-buffered).
    {
        std::vector<tfdtd_march> grids = {march(40, 30, 70, TFDTD_F32, 400), march(20, 18, 16, TFDTD_F32, 500)};
        const int n[2][3] = {{40, 30, 70}, {20, 18, 16}};
        for (int k = 0; k < 2; ++k) {
            grids[k].tile_y = 12;
            grids[k].chunk_x = 12;
            grids[k].h = family(n[k], 3, true, 450 + 100 * k);   // the two-step pass double-buffers it
        }
        const int count = (int)grids.size();
        const size_t bytes = tfdtd_batch_march_bytes(count, TFDTD_F32);
        std::vector<int32_t> prefix(3 * (count + 1));
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march2_pack(grids.data(), count, t, b, prefix.data()); }) == TFDTD_OK,
               "two-step pack of two grids");
        std::vector<tfdtd_march> bad = grids;
        bad[0].h.psi_out = const_cast<void*>(bad[0].h.psi_in);
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march2_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "an in-place H memory is refused by the two-step pass");
        bad = grids;
        bad[1].dtype = TFDTD_F64;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march2_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "float64 is refused by the two-step pass");
        std::vector<tfdtd_march> observed = grids;
        observed[0].snap_e = address(920);
        observed[0].snap_h = address(930);
        observed[0].snap_count = 2;
        const int box[2][2][3] = {{{3, 0, 0}, {5, 30, 70}}, {{30, 10, 5}, {33, 20, 60}}};
        for (int k = 0; k < 2; ++k)
            for (int ax = 0; ax < 3; ++ax) { observed[0].snap_lo[k][ax] = box[k][0][ax]; observed[0].snap_hi[k][ax] = box[k][1][ax]; }
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march2_pack(observed.data(), count, t, b, prefix.data()); }) == TFDTD_OK,
               "two-step pack with a step n+1 snapshot of two boxes");
        bad = observed;
        bad[0].snap_hi[1][0] = 41;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march2_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "a snapshot past the last plane is refused");
        bad = observed;
        bad[0].snap_lo[1][0] = 4;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march2_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "overlapping snapshot boxes are refused");
        bad = observed;
        bad[0].snap_hi[0][2] = 71;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march2_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "a snapshot box past the grid's z extent is refused");
        bad = observed;
        bad[0].snap_count = 5;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march2_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "more than four snapshot boxes are refused");
        bad = observed;
        bad[0].snap_h = nullptr;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_march2_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "a snapshot without its H buffer is refused");
    }
    // Split half-step tables.
    {
        std::vector<tfdtd_half_step> steps = {half_step(12, 10, 8, 600), half_step(30, 7, 5, 700), half_step(9, 9, 9, 800)};
        const int count = (int)steps.size();
        const size_t bytes = tfdtd_batch_table_bytes(count, TFDTD_F32);
        std::vector<int32_t> prefix(count + 1);
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_pack(steps.data(), count, t, b, prefix.data()); }) == TFDTD_OK,
               "half-step pack of three grids");
        expect(prefix[count] > prefix[0], "the half-step prefix grows");
        std::vector<tfdtd_half_step> bad = steps;
        bad[1].nseg = TFDTD_MAX_SEGMENTS + 1;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "too many CPML segments are refused");
        bad = steps;
        bad[2].seg[0].hi = 10;
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_batch_pack(bad.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "a segment past the axis is refused");
        expect(pack_exact(bytes / 2, [&](void* t, size_t b) { return tfdtd_batch_pack(steps.data(), count, t, b, prefix.data()); }) != TFDTD_OK,
               "a short half-step table is refused");
    }
    // State-norm tables.
    {
        std::vector<tfdtd_norm_grid> grids(2);
        std::memset(grids.data(), 0, grids.size() * sizeof grids[0]);
        const int dims[2][3] = {{40, 30, 20}, {7, 300, 9}};
        for (int g = 0; g < 2; ++g) {
            grids[g].dtype = TFDTD_F32;
            grids[g].nx = dims[g][0]; grids[g].ny = dims[g][1]; grids[g].nz = dims[g][2];
            grids[g].E = address(900 + 4 * g); grids[g].H = address(901 + 4 * g);
            grids[g].inverse_permittivity = address(902 + 4 * g);
            grids[g].eps_stride = 1;
            grids[g].psi[0] = address(903 + 4 * g);
            grids[g].psi_count[0] = 12345 + g;
        }
        const int64_t tasks = tfdtd_norm_tasks(grids.data(), 2);
        const size_t bytes = tfdtd_norm_table_bytes(grids.data(), 2);
        expect(tasks > 0 && bytes > 0, "norm tasks and table size");
        expect(pack_exact(bytes, [&](void* t, size_t b) { return tfdtd_norm_pack(grids.data(), 2, t, b); }) == TFDTD_OK,
               "norm pack of two grids");
        expect(pack_exact(bytes - 8, [&](void* t, size_t b) { return tfdtd_norm_pack(grids.data(), 2, t, b); }) != TFDTD_OK,
               "a short norm table is refused");
    }
    std::printf("%d checks, %d failures\n", checks, failures);
    return failures ? 1 : 0;
}
