// C ABI of the batched apply (gpaw_batch_*): groups of (operator, bands,
// output, Bloch phases) applied in one persistent launch per stencil range.
//
// Descriptors are written into a pinned host slot, copied to the device on
// the caller's stream and read by the kernel; a ring of slots with one
// event each lets the host fill the next call's descriptors while earlier
// launches still run.  Groups whose grid cannot be cut into tiles (tiny or
// very narrow periodic grids) and non-star stencils are applied by the
// single-operator kernels on the same stream.
#include <algorithm>
#include <vector>

#include "batch.h"
#include "capi_internal.h"

namespace gpaw {

hipError_t batch_launch_f64(const BatchGroup*, int, int, int, int, int, hipStream_t);
hipError_t batch_launch_c128(const BatchGroup*, int, int, int, int, int, hipStream_t);
hipError_t batch_launch_f32(const BatchGroup*, int, int, int, int, int, hipStream_t);
hipError_t batch_launch_c64(const BatchGroup*, int, int, int, int, int, hipStream_t);

hipError_t batch_launch(int dtype, const BatchGroup* d_groups, int ngroups, int ntiles, int R,
                        int TZ, int nt, hipStream_t st)
{
    switch (dtype) {
    case DT_F64: return batch_launch_f64(d_groups, ngroups, ntiles, R, TZ, nt, st);
    case DT_C128: return batch_launch_c128(d_groups, ngroups, ntiles, R, TZ, nt, st);
    case DT_F32: return batch_launch_f32(d_groups, ngroups, ntiles, R, TZ, nt, st);
    case DT_C64: return batch_launch_c64(d_groups, ngroups, ntiles, R, TZ, nt, st);
    }
    return hipErrorInvalidValue;
}

// Block size of the batch kernel for range R and storage type dtype.
static int batch_block(int dtype, int R)
{
    const bool cplx = dtype == DT_C128 || dtype == DT_C64;
    return (cplx && R <= 3) ? TILE_NT_WIDE : TILE_NT;
}

}  // namespace gpaw

using namespace gpaw;

namespace {
constexpr int SLOTS = 4;
// Planes per tile of a group: the whole axis up to 64 planes (each tile
// re-reads 2R queue planes, and a full periodic period reads none twice
// but the wrap), longer axes in equal runs of at most 64.  Tiles of
// different groups then cost within a small factor of each other, and the
// persistent grid takes the next tile as it frees up.
int batch_q(const Geom& g)
{
    const int n0 = g.n[0];
    if (n0 <= 64) return n0;
    const int runs = (n0 + 63) / 64;
    return (n0 + runs - 1) / runs;
}
}

struct gpaw_batch {
    int max_groups = 0;
    BatchGroup* h[SLOTS] = {};
    BatchGroup* d[SLOTS] = {};
    hipEvent_t ev[SLOTS] = {};
    bool used[SLOTS] = {};
    int next = 0;
};

extern "C" void gpaw_batch_free(gpaw_batch* b)
{
    if (!b) return;
    for (int s = 0; s < SLOTS; s++) {
        if (b->used[s]) hipEventSynchronize(b->ev[s]);
        if (b->ev[s]) hipEventDestroy(b->ev[s]);
        if (b->h[s]) hipHostFree(b->h[s]);
        if (b->d[s]) hipFree(b->d[s]);
    }
    delete b;
}

extern "C" int gpaw_batch_new(int max_groups, gpaw_batch** out)
{
    if (!out) return gpaw_fail(GPAW_EINVAL, "batch: NULL argument");
    *out = nullptr;
    if (max_groups < 1 || max_groups > 65536)
        return gpaw_fail(GPAW_EINVAL, "batch: max_groups = %d", max_groups);
    gpaw_batch* b = new gpaw_batch();
    b->max_groups = max_groups;
    const size_t bytes = sizeof(BatchGroup) * (size_t)max_groups;
    hipError_t e = hipSuccess;
    for (int s = 0; s < SLOTS && e == hipSuccess; s++) {
        e = hipHostMalloc((void**)&b->h[s], bytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc((void**)&b->d[s], bytes);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev[s], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        gpaw_batch_free(b);
        return gpaw_check(e, "batch: buffers");
    }
    *out = b;
    return GPAW_OK;
}

extern "C" int gpaw_batch_apply(gpaw_batch* b, int ngroups, gpaw_operator* const* ops,
                                const void* const* ins, void* const* outs, const long* nins,
                                const double* phases, int dtype, void* stream)
{
    if (!b) return gpaw_fail(GPAW_EINVAL, "batch: NULL handle");
    if (ngroups < 0 || ngroups > b->max_groups)
        return gpaw_fail(GPAW_EINVAL, "batch: %d groups, the batch holds %d", ngroups,
                         b->max_groups);
    if (ngroups == 0) return GPAW_OK;
    if (!ops || !ins || !outs || !nins) return gpaw_fail(GPAW_EINVAL, "batch: NULL array");
    hipStream_t st = (hipStream_t)stream;
    const int esize = dtype_size(dtype);
    // validate every group before any work is queued
    for (int i = 0; i < ngroups; i++) {
        if (!ops[i]) return gpaw_fail(GPAW_EINVAL, "batch: group %d has no operator", i);
        int rc = gpaw_check_dtype(ops[i]->real, dtype);
        if (rc) return rc;
        if (nins[i] < 0 || nins[i] > 65535)
            return gpaw_fail(GPAW_EINVAL, "batch: group %d nin = %ld", i, nins[i]);
        if (nins[i] && (!ins[i] || !outs[i]))
            return gpaw_fail(GPAW_EINVAL, "batch: group %d has a NULL array", i);
        if (nins[i] && ins[i] == outs[i])
            return gpaw_fail(GPAW_EINVAL, "batch: group %d input and output must differ", i);
    }
    // every group's tiles run concurrently: no output may overlap any input
    // or another output (inputs may overlap each other)
    {
        struct Span {
            uintptr_t lo, hi;
            int out, group;
        };
        std::vector<Span> sp;
        sp.reserve(2 * (size_t)ngroups);
        for (int i = 0; i < ngroups; i++) {
            if (!nins[i]) continue;
            const uintptr_t bytes = (uintptr_t)nins[i] * (uintptr_t)ops[i]->g.n[0] *
                                    (uintptr_t)ops[i]->g.n[1] * (uintptr_t)ops[i]->g.n[2] *
                                    (uintptr_t)esize;
            sp.push_back({(uintptr_t)ins[i], (uintptr_t)ins[i] + bytes, 0, i});
            sp.push_back({(uintptr_t)outs[i], (uintptr_t)outs[i] + bytes, 1, i});
        }
        std::sort(sp.begin(), sp.end(), [](const Span& a, const Span& b) { return a.lo < b.lo; });
        uintptr_t end_all = 0, end_out = 0;
        for (const Span& x : sp) {
            if (x.lo < end_out || (x.out && x.lo < end_all))
                return gpaw_fail(GPAW_EINVAL,
                                 "batch: group %d's %s overlaps another group's output or its "
                                 "input (the groups run concurrently)",
                                 x.group, x.out ? "output" : "input");
            end_all = std::max(end_all, x.hi);
            if (x.out) end_out = std::max(end_out, x.hi);
        }
    }
    // one slot of descriptors; wait until the launch that last used it ended
    const int s = b->next;
    b->next = (b->next + 1) % SLOTS;
    if (b->used[s]) {
        hipError_t e = hipEventSynchronize(b->ev[s]);
        if (e != hipSuccess) return gpaw_check(e, "batch: slot reuse");
    }
    BatchGroup* h = b->h[s];
    // per stencil range: tile-able groups first, the rest applied singly
    int count[GPAW_MAXR + 1] = {0};
    int tz_vote[GPAW_MAXR + 1] = {0};
    for (int i = 0; i < ngroups; i++) {
        const int R = ops[i]->st.star_r;
        if (R >= 1 && R <= 4 && nins[i]) tz_vote[R] += tile_tz(ops[i]->g.n[2]) == 64 ? 1 : -1;
    }
    int n = 0;
    int start[GPAW_MAXR + 2] = {0};
    int tiles[GPAW_MAXR + 1] = {0};
    int nts[GPAW_MAXR + 1] = {0};
    std::vector<char> fallback(ngroups, 0);
    for (int R = 1; R <= 4; R++) {
        start[R] = n;
        const int TZ = tz_vote[R] >= 0 ? 64 : 32;
        // wide blocks for complex data when every group of the range can be
        // cut into them (one launch has one block size), else 256
        int nt = batch_block(dtype, R);
        for (int i = 0; i < ngroups && nt != TILE_NT; i++) {
            TileShape sh;
            if (ops[i]->st.star_r == R && nins[i] &&
                !tile_geometry(ops[i]->g, R, esize, (int)nins[i], TZ, batch_q(ops[i]->g), sh,
                               nt) &&
                tile_geometry(ops[i]->g, R, esize, (int)nins[i], TZ, batch_q(ops[i]->g), sh,
                              TILE_NT))
                nt = TILE_NT;
        }
        nts[R] = nt;
        for (int i = 0; i < ngroups; i++) {
            gpaw_operator* op = ops[i];
            if (op->st.star_r != R || !nins[i]) continue;
            TileShape sh;
            if (!tile_geometry(op->g, R, esize, (int)nins[i], TZ, batch_q(op->g), sh, nt) ||
                (long)tiles[R] + sh.nblocks > 0x7fffffffL) {
                fallback[i] = 1;
                continue;
            }
            BatchGroup& gr = h[n++];
            gr.in = ins[i];
            gr.out = outs[i];
            gr.g = op->g;
            gpaw_phases(phases ? phases + 12 * i : nullptr, gr.ph);
            for (int c = 0; c < 6 * R + 1; c++) gr.sc.c[c] = op->st.coefs[c];
            gr.sh = sh;
            gr.ng = (i64)op->g.n[0] * op->g.n[1] * op->g.n[2];
            gr.first = tiles[R];
            gr.pad = 0;
            tiles[R] += sh.nblocks;
        }
        count[R] = n - start[R];
    }
    hipError_t e = hipSuccess;
    if (n) {
        e = hipMemcpyAsync(b->d[s], h, sizeof(BatchGroup) * n, hipMemcpyHostToDevice, st);
        for (int R = 1; R <= 4 && e == hipSuccess; R++) {
            if (!count[R]) continue;
            const int TZ = tz_vote[R] >= 0 ? 64 : 32;
            e = batch_launch(dtype, b->d[s] + start[R], count[R], tiles[R], R, TZ, nts[R], st);
        }
        // the slot is busy until the copy and the launches reading it are done
        const hipError_t er = hipEventRecord(b->ev[s], st);
        b->used[s] = (er == hipSuccess);
        if (e == hipSuccess) e = er;
    }
    for (int i = 0; i < ngroups && e == hipSuccess; i++) {
        gpaw_operator* op = ops[i];
        const int R = op->st.star_r;
        if (!nins[i] || (R >= 1 && R <= 4 && !fallback[i])) continue;
        Phases ph;
        gpaw_phases(phases ? phases + 12 * i : nullptr, ph);
        e = launch_stencil(op->st, dtype, ST_APPLY, ins[i], outs[i], nullptr, op->g, ph,
                           (int)nins[i], 0.0, nullptr, nullptr, true, st);
    }
    return gpaw_check(e, "batch apply");
}
