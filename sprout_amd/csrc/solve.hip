// solve.hip — the pressure solve on gfx950: conjugate gradients on two right-hand sides
// at once (pressure * dt and the density-correction potential), preconditioned by an
// aggregation multigrid V-cycle (mod.rs:2929-3010; mg_*.comp, cg_*.comp).
//
// Layout: every level of the hierarchy back to back in mga (float4 couplings L, R, B, T),
// mgd (diagonal), mgx / mgz / mgr (float2 per cell: pre-smoothed iterate, finished
// iterate, right-hand side), as the reference lays them out (mg.glsl:1-37). Kernels skip the
// 16 x 16 tiles without unknowns (ptile).
//
// Launches per CG iteration: the scalar step that ends the previous iteration and the one
// after cg_apply (one work-group each), cg_apply, the update fused with the way down on
// level 0, the coarse levels two per launch on the way down and up, the tail of the
// hierarchy in one work-group, and the way up on level 0: ten at 1024 x 576, where the
// reference records thirteen dispatches and two one-group reductions.
//
// Precision: cell arithmetic is FP32 in the reference's operation order; the scalar
// reductions (r.z, r.r, p.q, b.b) are FP64 — the product of two floats is exact in double
// and the sums are formed per work-group and then in a fixed order, so the scalars do not
// depend on scheduling. With mg_half the V-cycle's storage is narrower (MG<true> below);
// its arithmetic stays FP32, and the level-0 residual b - A x and the operator A p are
// evaluated in compensated FP32 (Ff below): the FP32 sums of those five-term stencils
// cancel (the pressure is near hydrostatic, |A| |x| ~ 1e4 |b|), and their rounding, not the
// tolerance, is what bounds the FP32 solver's error against the exact solution (1e-5 at
// 1024 x 576; 1e-7 compensated).
//
// Convergence without host round trips: the scalar state lives in device memory (sp_solve);
// every kernel of an iteration reads its `done` flags at entry and returns at once when
// both potentials have converged, so a batch can carry a fixed number of recorded iterations.
#include "dev.h"

namespace {

constexpr int NB_X[4] = {-1, 1, 0, 0};
constexpr int NB_Y[4] = {0, 0, -1, 1};

struct Lv {
    int off, w, h, toff;
};

SP_DEV Lv level(const Ctx& c, int l) {
    int4 L = c.lev[l];
    return Lv{L.x, L.y, L.z, L.w};
}
SP_DEV bool inside(const Lv& L, int x, int y) { return x >= 0 && y >= 0 && x < L.w && y < L.h; }
SP_DEV int at(const Lv& L, int x, int y) { return L.off + y * L.w + x; }
SP_DEV bool unknown(const Ctx& c, const Lv& L, int x, int y) { return inside(L, x, y) && c.mgd[at(L, x, y)] > 0.0f; }

// A sum of products carried as an unevaluated pair s + c of floats: each product a b is
// split exactly into its rounded value and error by an FMA (TwoProduct), each addition's
// rounding error is recovered by Knuth's TwoSum, and the errors are added into c. The
// result is as accurate as if the sum were formed in twice the working precision and then
// rounded (Ogita, Rump and Oishi, "Accurate sum and dot product", SIAM J. Sci. Comput. 26,
// 2005, algorithm Dot2). Needs -ffp-contract=off, which the library is built with.
struct Ff {
    float s, c;
};
SP_DEV Ff ff_prod(float a, float b) {
    const float p = a * b;
    return Ff{p, __builtin_fmaf(a, b, -p)};
}
SP_DEV Ff ff_add_prod(Ff acc, float a, float b) {
    const float p = a * b, e = __builtin_fmaf(a, b, -p);
    const float s = acc.s + p, bb = s - acc.s, err = (acc.s - (s - bb)) + (p - bb);
    return Ff{s, acc.c + (err + e)};
}
SP_DEV float ff_value(Ff a) { return a.s + a.c; }

// Storage of the multigrid levels. MG<false>: FP32, the reference's layout. MG<true>
// (mg_half): every coarse level (1 and below) narrow: iterates (mgx, mgz) and right-hand
// sides (mgr) in BF16, two per 32-bit word; couplings and diagonals in FP16, scaled by 2^-l
// so that their growth with the level (about 2x per level under 2 x 2 aggregation) stays
// far from FP16's range. Level 0 stays FP32: its matrix is also the CG operator, its
// residual a CG vector, and its finished iterate z the preconditioned residual CG steps
// along (in BF16 that last one cost 1.8x the iterations at 1024 x 576). Narrow values sit at
// the start of each array's coarse region, which begins at cell N (level 1). Loads widen to
// FP32 exactly, so all arithmetic is FP32; stores round to nearest even.
#ifndef SP_MG_NARROW_ITERATES
#define SP_MG_NARROW_ITERATES 1
#endif
struct alignas(8) Half4 {
    _Float16 x, y, z, w;
};
SP_DEV uint32_t bf16_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
SP_DEV uint32_t pack_bf2(float2 v) { return bf16_bits(v.x) | (bf16_bits(v.y) << 16); }
SP_DEV float2 unpack_bf2(uint32_t w) { return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u)); }

template <bool H>
struct MG {
    SP_DEV static float D(const Ctx& c, int, int i) { return c.mgd[i]; }
    SP_DEV static float4 A(const Ctx& c, int, int i) { return c.mga[i]; }
    SP_DEV static void set_AD(const Ctx& c, int, int i, float4 a, float d) {
        c.mgd[i] = d;
        c.mga[i] = a;
    }
    SP_DEV static float2 X(const Ctx& c, int i) { return c.mgx[i]; }
    SP_DEV static void set_X(const Ctx& c, int i, float2 v) { c.mgx[i] = v; }
    SP_DEV static float2 Z(const Ctx& c, int i) { return c.mgz[i]; }
    SP_DEV static void set_Z(const Ctx& c, int i, float2 v) { c.mgz[i] = v; }
    // right-hand sides of the coarse levels (i >= N)
    SP_DEV static float2 R(const Ctx& c, int i) { return c.mgr[i]; }
    SP_DEV static void set_R(const Ctx& c, int i, float2 v) { c.mgr[i] = v; }
    // a coarse right-hand side or iterate as it reads back after storing
    SP_DEV static float2 stored_R(float2 v) { return v; }
    SP_DEV static float2 stored_Z(float2 v) { return v; }
};

template <>
struct MG<true> {
    SP_DEV static float D(const Ctx& c, int l, int i) {
        if (l == 0) return c.mgd[i];
        return ldexpf((float)reinterpret_cast<const _Float16*>(c.mgd + c.N)[i - c.N], l);
    }
    SP_DEV static float4 A(const Ctx& c, int l, int i) {
        if (l == 0) return c.mga[i];
        const Half4 h = reinterpret_cast<const Half4*>(c.mga + c.N)[i - c.N];
        return make_float4(ldexpf((float)h.x, l), ldexpf((float)h.y, l), ldexpf((float)h.z, l), ldexpf((float)h.w, l));
    }
    SP_DEV static void set_AD(const Ctx& c, int l, int i, float4 a, float d) {
        if (l == 0) {
            c.mgd[i] = d;
            c.mga[i] = a;
            return;
        }
        reinterpret_cast<_Float16*>(c.mgd + c.N)[i - c.N] = (_Float16)ldexpf(d, -l);
        Half4 h;
        h.x = (_Float16)ldexpf(a.x, -l);
        h.y = (_Float16)ldexpf(a.y, -l);
        h.z = (_Float16)ldexpf(a.z, -l);
        h.w = (_Float16)ldexpf(a.w, -l);
        reinterpret_cast<Half4*>(c.mga + c.N)[i - c.N] = h;
    }
    SP_DEV static float2 vec(const float2* a, const Ctx& c, int i) {
        if (i < c.N) return a[i];
        return unpack_bf2(reinterpret_cast<const uint32_t*>(a + c.N)[i - c.N]);
    }
    SP_DEV static void set_vec(float2* a, const Ctx& c, int i, float2 v) {
        if (i < c.N) a[i] = v;
        else reinterpret_cast<uint32_t*>(a + c.N)[i - c.N] = pack_bf2(v);
    }
#if SP_MG_NARROW_ITERATES
    SP_DEV static float2 X(const Ctx& c, int i) { return vec(c.mgx, c, i); }
    SP_DEV static void set_X(const Ctx& c, int i, float2 v) { set_vec(c.mgx, c, i, v); }
    SP_DEV static float2 Z(const Ctx& c, int i) { return vec(c.mgz, c, i); }
    SP_DEV static void set_Z(const Ctx& c, int i, float2 v) { set_vec(c.mgz, c, i, v); }
#else
    SP_DEV static float2 X(const Ctx& c, int i) { return c.mgx[i]; }
    SP_DEV static void set_X(const Ctx& c, int i, float2 v) { c.mgx[i] = v; }
    SP_DEV static float2 Z(const Ctx& c, int i) { return c.mgz[i]; }
    SP_DEV static void set_Z(const Ctx& c, int i, float2 v) { c.mgz[i] = v; }
#endif
    SP_DEV static float2 R(const Ctx& c, int i) {
        return unpack_bf2(reinterpret_cast<const uint32_t*>(c.mgr + c.N)[i - c.N]);
    }
    SP_DEV static void set_R(const Ctx& c, int i, float2 v) { reinterpret_cast<uint32_t*>(c.mgr + c.N)[i - c.N] = pack_bf2(v); }
    SP_DEV static float2 stored_R(float2 v) { return unpack_bf2(pack_bf2(v)); }
#if SP_MG_NARROW_ITERATES
    SP_DEV static float2 stored_Z(float2 v) { return unpack_bf2(pack_bf2(v)); }
#else
    SP_DEV static float2 stored_Z(float2 v) { return v; }
#endif
};

SP_DEV int tid2() { return threadIdx.y * 16 + threadIdx.x; }
SP_DEV int tile_slot() { return blockIdx.y * gridDim.x + blockIdx.x; }

// The tile (of a grid of them, gridDim.x x gridDim.y) a work-group of the solve works on.
// Work-groups go to the eight XCDs in turn, and each XCD has an L2 of its own, so
// consecutive work-groups would sit on different L2s and fetch their neighbours' rings
// again; here each XCD is handed a contiguous run of tiles (whole rows of them), so that the
// rings mostly come from its own L2. A bijection of the tiles: only the speed depends on how
// work-groups are dispatched.
SP_DEV int2 tile_xy() {
    const int nx = (int)gridDim.x, n = nx * (int)gridDim.y;
    const int b = (int)(blockIdx.y * gridDim.x + blockIdx.x);
    const int q = n >> 3, r = n & 7, xcd = b & 7, k = b >> 3;
    const int t = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + k;
    return make_int2(t % nx, t / nx);
}
SP_DEV int slot_of(int2 tq) { return tq.y * (int)gridDim.x + tq.x; }

SP_DEV void add4(double4& a, const double4& b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    a.w += b.w;
}

// FP64 sum of one double4 per work-item over a group of NT work-items in a fixed order: a
// butterfly across each wave64 (x + y = y + x exactly, so every lane holds the same sum),
// then the waves in order. Every work-item gets the total; `lds` (NT / 64 entries) may be
// reused once the call returns. Every work-item of the group must call it.
template <int NT = 256>
SP_DEV double4 group_sum4(double4 v, double4* lds) {
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o);
        v.y += __shfl_xor(v.y, o);
        v.z += __shfl_xor(v.z, o);
        v.w += __shfl_xor(v.w, o);
    }
    if ((t & 63) == 0) lds[t >> 6] = v;
    __syncthreads();
    double4 s = lds[0];
    // in order, a few waves' sums in flight at a time (all of them at once would hold
    // 8 registers per wave of the group)
#pragma unroll 3
    for (int w = 1; w < NT / 64; w++) add4(s, lds[w]);
    __syncthreads();
    return s;
}

SP_DEV void store_part(double4* p, double4 v) { *p = v; }

// Sum of parts[base .. base + n) by one group of NT work-items in a fixed order (work-item
// t adds entries t, t + NT, t + 2 NT ... in turn); each work-item issues its loads eight at a
// time before adding, so their latencies overlap (2304 partial sums at 1024 x 576: two
// rounds for 256 work-items).
template <int NT>
SP_DEV double4 total_of(const double4* parts, int base, int n, double4* lds) {
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    double4 s = make_double4(0, 0, 0, 0);
    for (int k0 = 0; k0 < n; k0 += 8 * NT) {
        double4 v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + j * NT + t;
            v[j] = k < n ? parts[base + k] : make_double4(0, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 8; j++) add4(s, v[j]);
    }
    return group_sum4<NT>(s, lds);
}
constexpr int REDUCE_NT = 256;  // work-items of the one-group reduction kernels
SP_DEV double4 total(const Ctx& c, int base, int n, double4* lds) { return total_of<REDUCE_NT>(c.part_sums, base, n, lds); }

}  // namespace

// ---- hierarchy set-up -------------------------------------------------------------------

// Galerkin coarse matrix of coarse cell (cx, cy) of level lf + 1 from fine level F = lf
// (mg_build.comp:43-65).
template <bool H>
SP_DEV void build_cell(const Ctx& c, const Lv& F, const Lv& C, int lf, int cx, int cy) {
    using M = MG<H>;
    float D = 0.0f;
    float4 A = make_float4(0, 0, 0, 0);
    for (int k = 0; k < 4; k++) {
        int x = 2 * cx + (k & 1), y = 2 * cy + (k >> 1);
        if (!(inside(F, x, y) && M::D(c, lf, at(F, x, y)) > 0.0f)) continue;
        int i = at(F, x, y);
        D += M::D(c, lf, i);
        float4 w = M::A(c, lf, i);
        for (int d = 0; d < 4; d++) {
            int nx = x + NB_X[d], ny = y + NB_Y[d];
            if (!(inside(F, nx, ny) && M::D(c, lf, at(F, nx, ny)) > 0.0f)) continue;
            float wd = f4c(w, d);
            if ((nx >> 1) == cx && (ny >> 1) == cy) D -= wd;
            else if (d == 0) A.x += wd;
            else if (d == 1) A.y += wd;
            else if (d == 2) A.z += wd;
            else A.w += wd;
        }
    }
    const int o = at(C, cx, cy);
    if (D > 1e-7f) M::set_AD(c, lf + 1, o, A, D);
    else M::set_AD(c, lf + 1, o, make_float4(0, 0, 0, 0), 0.0f);
}

// mg_build.comp:16-65: coarse level l from level l - 1, over coarse 16 x 16 tiles.
template <bool H>
__global__ __launch_bounds__(256) void k_mg_build(const TankBatch tb, int l) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ uint32_t s_state;
    const Lv F = level(c, l - 1), C = level(c, l);
    if (tid2() == 0) {
        int tfx = (F.w + 15) >> 4, tfy = (F.h + 15) >> 4;
        uint32_t on = 0;
        for (int k = 0; k < 4; k++) {
            int tx = 2 * (int)blockIdx.x + (k & 1), ty = 2 * (int)blockIdx.y + (k >> 1);
            if (tx < tfx && ty < tfy) on |= c.ptile[F.toff + ty * tfx + tx];
        }
        uint32_t slot = C.toff + tile_slot();
        uint32_t was = c.ptile[slot];
        c.ptile[slot] = on ? 1u : 0u;
        s_state = (on ? 1u : 0u) | (was ? 2u : 0u);
    }
    __syncthreads();
    uint32_t state = s_state;
    if (state == 0u) return;
    int cx = blockIdx.x * 16 + threadIdx.x, cy = blockIdx.y * 16 + threadIdx.y;
    if (!inside(C, cx, cy)) return;
    if ((state & 1u) == 0u) {
        MG<H>::set_AD(c, l, at(C, cx, cy), make_float4(0, 0, 0, 0), 0.0f);
        return;
    }
    build_cell<H>(c, F, C, l - 1, cx, cy);
}

// mg_tail.comp mode 1: the coarse matrices of every level below the tail, one group.
template <bool H>
__global__ __launch_bounds__(1024) void k_mg_tail_build(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    for (int l = c.tail; l < c.n_levels - 1; l++) {
        const Lv F = level(c, l), C = level(c, l + 1);
        for (int t = threadIdx.x; t < C.w * C.h; t += 1024) build_cell<H>(c, F, C, l, t % C.w, t / C.w);
        __threadfence_block();
        __syncthreads();
    }
}

// ---- Krylov kernels on level 0 ---------------------------------------------------------

// cg_scale.comp:14-27: rescale the warm start to this substep's length. With a second solve
// after the viscosity that warm-starts from its own part (mode 1), the first starts from
// P1 = P - P2 and P2 is kept; otherwise P2 is forgotten.
__global__ __launch_bounds__(256) void k_cg_scale(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    if (c.ptile[tile_slot()] == 0u) return;
    const Lv F = level(c, 0);
    int x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (!inside(F, x, y)) return;
    int i = at(F, x, y);
    if (c.mgd[i] > 0.0f) {
        float p = c.pres[i].x;
        if (c.P->v_p2warm) p -= c.p2keep[i];
        else c.p2keep[i] = 0.0f;
        c.pres[i].x = p * (c.S->dt / c.S->dt_prev);
    } else {
        c.pres[i] = make_float2(0.0f, 0.0f);
    }
}

// ---- the CG's scalar steps -------------------------------------------------------------
//
// The reference folds each scalar step into the last work-group of the kernel that made the
// partial sums (reduce_tail.glsl:28-57, behind an atomic arrival and a device-wide fence).
// Here a one-work-group kernel per step sums the partials in a fixed order after the kernel
// boundary has ordered them: on a part with eight L2 caches a device-scope release in every
// producing work-group costs more than a launch. (Letting every work-group of the consuming
// kernel sum all 2304 partials at 1024 x 576 instead moves about 20 MB through each L2 per
// step, which costs more again.) The scalar state of iteration k is kept in slot k & 1 of the
// solve record (sp_cg_state): the step that starts iteration k reads slot (k - 1) & 1.
//
// Partial sums, per level-0 block (nbp of them): [0, nbp) r.z (xy), r.r (zw) of the last
// level-0 up-sweep; [nbp, 2 nbp) b.b (xy) of cg_init; [2 nbp, 3 nbp) p.q (xy) of cg_apply.

namespace {

constexpr int T0 = 16;          // level-0 block of the Krylov kernels: TT x TT tiles of 16 x 16
constexpr int TT = T0 / 16;
constexpr int NT0 = T0 * T0;    // its work-items
constexpr int PART_RZ = 0, PART_BB = 1, PART_PQ = 2;

SP_DEV int tid0() { return threadIdx.y * T0 + threadIdx.x; }
SP_DEV bool any_on(const uint32_t* s_on) {
    uint32_t on = 0u;
    for (int k = 0; k < TT * TT; k++) on |= s_on[k];
    return on != 0u;
}
SP_DEV int slot0() { return slot_of(tile_xy()); }

SP_DEV bool st_done(const sp_cg_state& s) { return (s.done[0] != 0u && s.done[1] != 0u) || s.halt != 0u; }
// A solve is over for this tank, in the iteration of parity `par`, when both potentials
// converged or when it has done the iterations this tank records (its cap).
SP_DEV bool slot_done(const Ctx& c, int par) {
    const sp_cg_state& s = c.K->st[par];
    return (s.done[0] != 0u && s.done[1] != 0u) || s.halt != 0u;
}

// Whether 16 x 16 tile (tx, ty) of level L has unknowns (ptile); tiles outside do not.
SP_DEV bool tile_on(const Ctx& c, const Lv& L, int tx, int ty) {
    const int tw = (L.w + 15) >> 4, th = (L.h + 15) >> 4;
    return tx >= 0 && ty >= 0 && tx < tw && ty < th && c.ptile[L.toff + ty * tw + tx] != 0u;
}

// The start of a solve (reduce_tail.glsl:59-75, mode 0; mode 8 for the second solve after
// the viscosity, where the density-correction potential counts as solved): a = r.z (xy),
// r.r (zw) of the first V-cycle, b = b.b (xy).
SP_DEV sp_cg_state cg_start(const Ctx& c, const double4& a, const double4& b, bool second) {
    const sp_gparams& P = *c.P;
    const double tx = (double)P.tol[0] * (double)P.tol[0], ty = (double)P.tol[1] * (double)P.tol[1];
    sp_cg_state s = {};
    s.bb[0] = b.x;
    s.bb[1] = b.y;
    s.rz[0] = a.x;
    s.rz[1] = a.y;
    s.rr[0] = a.z;
    s.rr[1] = a.w;
    s.done[0] = a.z <= tx * b.x ? 1u : 0u;
    s.done[1] = (second || a.w <= ty * b.y) ? 1u : 0u;
    s.cap = second ? c.K->cap_p2 : c.K->cap_p;
    return s;
}

// The end of an iteration (reduce_tail.glsl:77-104, mode 2): from the state e entering it and
// a = r.z (xy), r.r (zw) of its up-sweep, the state entering the next one: the stop test
// r.r <= tol^2 b.b per potential, beta, and the halt at this tank's recorded iterations.
SP_DEV sp_cg_state cg_next(const Ctx& c, const sp_cg_state& e, const double4& a) {
    const sp_gparams& P = *c.P;
    const double tx = (double)P.tol[0] * (double)P.tol[0], ty = (double)P.tol[1] * (double)P.tol[1];
    sp_cg_state s = e;
    const uint32_t was0 = e.done[0], was1 = e.done[1];
    s.iters[0] += 1u - was0;
    s.iters[1] += 1u - was1;
    const uint32_t d0 = was0 | (a.z <= tx * e.bb[0] ? 1u : 0u);
    const uint32_t d1 = was1 | (a.w <= ty * e.bb[1] ? 1u : 0u);
    s.beta[0] = (d0 == 0u && e.rz[0] > 0.0) ? a.x / e.rz[0] : 0.0;
    s.beta[1] = (d1 == 0u && e.rz[1] > 0.0) ? a.y / e.rz[1] : 0.0;
    s.rz[0] = a.x;
    s.rz[1] = a.y;
    if (!was0) s.rr[0] = a.z;
    if (!was1) s.rr[1] = a.w;
    s.done[0] = d0;
    s.done[1] = d1;
    s.sweeps += 1u;
    if (s.cap != 0u && s.sweeps >= s.cap && (d0 == 0u || d1 == 0u)) s.halt = 1u;
    return s;
}

// Index of global cell (x, y) in a staging array of side S holding the region that starts r
// cells before (x0, y0).
SP_DEV int sx_(int x, int y, int x0, int y0, int r, int S) { return (y - y0 + r) * S + (x - x0 + r); }

SP_DEV float2 direction_of(float2 z, float2 po, float bx, float by) {
    if (bx == 0.0f && by == 0.0f) return z;
    return make_float2(bx != 0.0f ? z.x + bx * po.x : z.x, by != 0.0f ? z.y + by * po.y : z.y);
}

// ---- one level of the V-cycle on a T x T block: the arithmetic of mg_down.comp:36-69 and
// mg_up.comp:26-75, operands from LDS ------------------------------------------------------

// Down, from zero (mg_down.comp:36-60): z on the block plus one ring from the right-hand side
// r and diagonal D on two rings (S2 = T + 4) and the couplings A on one (S1 = T + 2).
template <int NT>
SP_DEV void down_z(int x0, int y0, int S1, int S2, const float* sd, const float2* sr, const float4* sa, float2* sz) {
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    for (int k = t; k < S1 * S1; k += NT) {
        const int gx = x0 - 1 + k % S1, gy = y0 - 1 + k / S1;
        const int s2 = sx_(gx, gy, x0, y0, 2, S2);
        const float D = sd[s2];
        float2 z = make_float2(0.0f, 0.0f);
        if (D > 0.0f) {
            const float2 r = sr[s2];
            if (((gx + gy) & 1) == 0) {
                z = make_float2(r.x / D, r.y / D);
            } else {
                const float4 w = sa[k];
                float dj[4];
                float2 rj[4];
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    const int sn = sx_(gx + NB_X[d], gy + NB_Y[d], x0, y0, 2, S2);
                    dj[d] = sd[sn];
                    rj[d] = sr[sn];
                }
                float2 sum = r;
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    const float wd = f4c(w, d);
                    const float2 cand = make_float2(sum.x + wd * rj[d].x / dj[d], sum.y + wd * rj[d].y / dj[d]);
                    sum = dj[d] > 0.0f ? cand : sum;
                }
                z = make_float2(sum.x / D, sum.y / D);
            }
        }
        sz[k] = z;
    }
}

// The restriction of the residual left by down_z to coarse cell (cx, cy) (mg_down.comp:
// 61-69): its red children's neighbours' z weighted by the couplings. (qx0, qy0): the fine
// cell at index 0 of the ring-1 arrays (sz, sa); of the ring-2 array sd: (qx0 - 1, qy0 - 1).
SP_DEV float2 restrict_red(int cx, int cy, int qx0, int qy0, int S1, int S2, const float* sd, const float4* sa, const float2* sz) {
    float2 sum = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int qx = 2 * cx + k, qy = 2 * cy + k;
        const bool okq = sd[(qy - qy0 + 1) * S2 + (qx - qx0 + 1)] > 0.0f;
        const float4 w = sa[(qy - qy0) * S1 + (qx - qx0)];
        bool ok[4];
        float2 zn[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const int nx = qx + NB_X[d], ny = qy + NB_Y[d];
            ok[d] = okq && sd[(ny - qy0 + 1) * S2 + (nx - qx0 + 1)] > 0.0f;
            zn[d] = sz[(ny - qy0) * S1 + (nx - qx0)];
        }
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const float wd = f4c(w, d);
            const float2 cand = make_float2(sum.x + wd * zn[d].x, sum.y + wd * zn[d].y);
            sum = ok[d] ? cand : sum;
        }
    }
    return sum;
}

// Up (mg_up.comp:26-36): the black cells' new values from the corrected iterate x + w z_parent
// of their red neighbours, for the cells of the block plus `ring` rings. Region of the
// arrays: sd, sx from (x0 - rx, y0 - rx), side SX; sa, sr from (x0 - ring, y0 - ring), side SR
// (= the output's region); the parents' z from coarse cell (pcx, pcy), side SP.
template <int NT>
SP_DEV void up_black(int x0, int y0, int ring, int rx, int SX, int SR, int pcx, int pcy, int SP, const float* sd, const float2* sx,
                     const float4* sa, const float2* sr, const float2* szc, float2* sb) {
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    for (int k = t; k < SR * SR; k += NT) {
        const int gx = x0 - ring + k % SR, gy = y0 - ring + k / SR;
        float2 z = make_float2(0.0f, 0.0f);
        const float D = sd[sx_(gx, gy, x0, y0, rx, SX)];
        if (D > 0.0f && ((gx + gy) & 1) != 0) {
            const float4 w = sa[k];
            bool ok[4];
            float2 xm[4], zc[4];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int mx = gx + NB_X[d], my = gy + NB_Y[d];
                const int sm = sx_(mx, my, x0, y0, rx, SX);
                ok[d] = sd[sm] > 0.0f;
                xm[d] = sx[sm];
                zc[d] = szc[((my >> 1) - pcy) * SP + ((mx >> 1) - pcx)];
            }
            float2 sum = sr[k];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const float wd = f4c(w, d);
                const float2 cand = make_float2(sum.x + wd * (xm[d].x + SP_MG_OMEGA * zc[d].x),
                                                sum.y + wd * (xm[d].y + SP_MG_OMEGA * zc[d].y));
                sum = ok[d] ? cand : sum;
            }
            z = make_float2(sum.x / D, sum.y / D);
        }
        sb[k] = z;
    }
}

// Up (mg_up.comp:37-75): z of cell (x, y) after the red half-sweep over the black values sb
// (region from (x0 - rb, y0 - rb), side SB); sd from (x0 - rx, y0 - rx), side SX; sa, sr at
// index k of a region from (x0 - rb, y0 - rb) of side SB as well.
SP_DEV float2 up_red(int x, int y, int x0, int y0, int rb, int SB, int rx, int SX, const float* sd, const float4* sa, const float2* sr,
                     const float2* sb, float2* rv_out) {
    const float D = sd[sx_(x, y, x0, y0, rx, SX)];
    const int s1 = sx_(x, y, x0, y0, rb, SB);
    float2 z = make_float2(0.0f, 0.0f);
    *rv_out = make_float2(0.0f, 0.0f);
    if (!(D > 0.0f)) return z;
    const float2 rv = sr[s1];
    *rv_out = rv;
    if (((x + y) & 1) == 0) {
        const float4 w = sa[s1];
        bool ok[4];
        float2 zn[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const int nx = x + NB_X[d], ny = y + NB_Y[d];
            ok[d] = sd[sx_(nx, ny, x0, y0, rx, SX)] > 0.0f;
            zn[d] = sb[sx_(nx, ny, x0, y0, rb, SB)];
        }
        float2 sum = rv;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const float wd = f4c(w, d);
            const float2 cand = make_float2(sum.x + wd * zn[d].x, sum.y + wd * zn[d].y);
            sum = ok[d] ? cand : sum;
        }
        z = make_float2(sum.x / D, sum.y / D);
    } else {
        z = sb[s1];
    }
    return z;
}

}  // namespace

// ---- Krylov kernels on level 0 -----------------------------------------------------------

// cg_init.comp: r = b - A x and the partial sums of b.b (H: r compensated).
template <bool H>
__global__ __launch_bounds__(NT0) void k_cg_init(const TankBatch tb, Lv F, int nbp) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ double4 lds[NT0 / 64];
    __shared__ uint32_t s_on[TT * TT];
    const int t = tid0();
    const int2 tq = tile_xy();
    const int x0 = tq.x * T0, y0 = tq.y * T0;
    if (t < TT * TT) s_on[t] = tile_on(c, F, (x0 >> 4) + t % TT, (y0 >> 4) + t / TT) ? 1u : 0u;
    __syncthreads();
    if (!any_on(s_on)) {
        if (t == 0) store_part(&c.part_sums[PART_BB * nbp + slot0()], make_double4(0, 0, 0, 0));
        return;
    }
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    double4 acc = make_double4(0, 0, 0, 0);
    if (inside(F, x, y) && s_on[(threadIdx.y >> 4) * TT + (threadIdx.x >> 4)]) {
        const int i = at(F, x, y);
        const float D = c.mgd[i];
        if (D > 0.0f) {
            const float2 xv = c.pres[i];
            const float4 w = c.mga[i];
            const float4 b = c.prhs[i];
            if (H) {
                Ff rx = ff_add_prod(Ff{b.x, 0.0f}, -D, xv.x), ry = ff_add_prod(Ff{b.y, 0.0f}, -D, xv.y);
                for (int d = 0; d < 4; d++) {
                    const int nx = x + NB_X[d], ny = y + NB_Y[d];
                    if (unknown(c, F, nx, ny)) {
                        const float wd = f4c(w, d);
                        const float2 pn = c.pres[at(F, nx, ny)];
                        rx = ff_add_prod(rx, wd, pn.x);
                        ry = ff_add_prod(ry, wd, pn.y);
                    }
                }
                c.mgr[i] = make_float2(ff_value(rx), ff_value(ry));
            } else {
                float2 ax = make_float2(D * xv.x, D * xv.y);
                for (int d = 0; d < 4; d++) {
                    const int nx = x + NB_X[d], ny = y + NB_Y[d];
                    if (unknown(c, F, nx, ny)) {
                        const float wd = f4c(w, d);
                        const float2 pn = c.pres[at(F, nx, ny)];
                        ax.x -= wd * pn.x;
                        ax.y -= wd * pn.y;
                    }
                }
                c.mgr[i] = make_float2(b.x - ax.x, b.y - ax.y);
            }
            acc.x = (double)b.x * (double)b.x;
            acc.y = (double)b.y * (double)b.y;
        } else {
            c.mgr[i] = make_float2(0.0f, 0.0f);
        }
    }
    const double4 s = group_sum4<NT0>(acc, lds);
    if (t == 0) store_part(&c.part_sums[PART_BB * nbp + slot0()], s);
}

// The scalar step that starts iteration k (par = k & 1), one work-group per tank: the
// solve's start (reduce_tail.glsl:59-75, mode 0; mode 8 for the second solve) for k = 0 from
// b.b and the first V-cycle's r.z, r.r, otherwise the end of iteration k - 1 (mode 2) from its
// r.z, r.r partial sums; the state goes into slot par. A solve that is over passes its state on.
__global__ __launch_bounds__(REDUCE_NT) void k_cg_step(const TankBatch tb, int nbp, int par, int first, int second) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ double4 lds[REDUCE_NT / 64];
    const sp_cg_state* P = &c.K->st[par ^ 1];
    if (!first && st_done(*P)) {
        if (threadIdx.x == 0) c.K->st[par] = *P;
        return;
    }
    const double4 a = total_of<REDUCE_NT>(c.part_sums, PART_RZ * nbp, nbp, lds);
    double4 b = make_double4(0, 0, 0, 0);
    if (first) b = total_of<REDUCE_NT>(c.part_sums, PART_BB * nbp, nbp, lds);
    if (threadIdx.x != 0) return;
    if (first) {
        c.K->st[par] = cg_start(c, a, b, second != 0);
        c.K->solves += 1u;
    } else {
        c.K->st[par] = cg_next(c, *P, a);
    }
}

// cg_apply.comp:23-51: p = z + beta p_old, q = A p (H: compensated) and the partial sums of
// p.q. par = k & 1; the direction alternates between cgp and cgp2 with it.
template <bool H>
__global__ __launch_bounds__(NT0) void k_cg_apply(const TankBatch tb, Lv F, int nbp, int par) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const float2* __restrict__ p_old = par ? c.cgp2 : c.cgp;
    float2* __restrict__ p_new = par ? c.cgp : c.cgp2;
    constexpr int S1 = T0 + 2;
    __shared__ double4 lds[NT0 / 64];
    __shared__ float sd[S1 * S1];
    __shared__ float2 sp[S1 * S1];
    __shared__ uint32_t s_on[TT * TT];
    const int t = tid0();
    const int2 tq = tile_xy();
    const int x0 = tq.x * T0, y0 = tq.y * T0;
    if (t < TT * TT) s_on[t] = tile_on(c, F, (x0 >> 4) + t % TT, (y0 >> 4) + t / TT) ? 1u : 0u;
    const sp_cg_state* E = &c.K->st[par];
    if (st_done(*E)) return;
    __syncthreads();
    if (!any_on(s_on)) {
        if (t == 0) store_part(&c.part_sums[PART_PQ * nbp + slot0()], make_double4(0, 0, 0, 0));
        return;
    }
    const float bx = (float)E->beta[0], by = (float)E->beta[1];
    for (int k = t; k < S1 * S1; k += NT0) {
        const int gx = x0 - 1 + k % S1, gy = y0 - 1 + k / S1;
        float d = 0.0f;
        float2 pv = make_float2(0.0f, 0.0f);
        if (inside(F, gx, gy)) {
            const int i = at(F, gx, gy);
            d = c.mgd[i];
            if (d > 0.0f) {
                const float2 po = (bx == 0.0f && by == 0.0f) ? make_float2(0.0f, 0.0f) : p_old[i];
                pv = direction_of(MG<H>::Z(c, i), po, bx, by);
            }
        }
        sd[k] = d;
        sp[k] = pv;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    double4 acc = make_double4(0, 0, 0, 0);
    if (inside(F, x, y) && s_on[(threadIdx.y >> 4) * TT + (threadIdx.x >> 4)]) {
        const int i = at(F, x, y);
        const int si = sx_(x, y, x0, y0, 1, S1);
        const float D = sd[si];
        if (D > 0.0f) {
            const float2 p = sp[si];
            const float4 w = c.mga[i];
            float2 q;
            if (H) {
                Ff qx = ff_prod(D, p.x), qy = ff_prod(D, p.y);
                for (int d = 0; d < 4; d++) {
                    const int sn = sx_(x + NB_X[d], y + NB_Y[d], x0, y0, 1, S1);
                    if (sd[sn] > 0.0f) {
                        const float wd = f4c(w, d);
                        const float2 pn = sp[sn];
                        qx = ff_add_prod(qx, -wd, pn.x);
                        qy = ff_add_prod(qy, -wd, pn.y);
                    }
                }
                q = make_float2(ff_value(qx), ff_value(qy));
            } else {
                q = make_float2(D * p.x, D * p.y);
                for (int d = 0; d < 4; d++) {
                    const int sn = sx_(x + NB_X[d], y + NB_Y[d], x0, y0, 1, S1);
                    if (sd[sn] > 0.0f) {
                        const float wd = f4c(w, d);
                        const float2 pn = sp[sn];
                        q.x -= wd * pn.x;
                        q.y -= wd * pn.y;
                    }
                }
            }
            p_new[i] = p;
            c.cgq[i] = q;
            acc.x = (double)p.x * (double)q.x;
            acc.y = (double)p.y * (double)q.y;
        } else {
            p_new[i] = make_float2(0.0f, 0.0f);
        }
    }
    const double4 s = group_sum4<NT0>(acc, lds);
    if (t == 0) store_part(&c.part_sums[PART_PQ * nbp + slot0()], s);
}

// cg_apply's scalar step (reduce_tail.glsl mode 1), one work-group per tank: alpha = rz / p.q
// for each potential still running, into the iteration's state.
__global__ __launch_bounds__(REDUCE_NT) void k_cg_alpha(const TankBatch tb, int nbp, int par) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ double4 lds[REDUCE_NT / 64];
    sp_cg_state* E = &c.K->st[par];
    if (st_done(*E)) return;
    const double4 pq = total_of<REDUCE_NT>(c.part_sums, PART_PQ * nbp, nbp, lds);
    if (threadIdx.x != 0) return;
    E->alpha[0] = (E->done[0] == 0u && pq.x > 0.0) ? E->rz[0] / pq.x : 0.0;
    E->alpha[1] = (E->done[1] == 0u && pq.y > 0.0) ? E->rz[1] / pq.y : 0.0;
}

// cg_update.comp fused with the way down on level 0: x += alpha p, r_out = r_in - alpha q (residual buffers alternate with
// the iteration's parity, so a block may read its neighbours' old residual while they write
// their new one), then mg_down on level 0 from r_out: the pre-smoothed iterate and the
// level-1 right-hand side.
template <bool H>
__global__ __launch_bounds__(NT0) void k_update_down0(const TankBatch tb, Lv F, Lv C, int par) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const float2* __restrict__ r_in = par ? c.r2 : c.mgr;
    float2* __restrict__ r_out = par ? c.mgr : c.r2;
    const float2* __restrict__ p_new = par ? c.cgp : c.cgp2;
    constexpr int S1 = T0 + 2, S2 = T0 + 4;
    __shared__ float sd[S2 * S2];
    __shared__ float2 sr[S2 * S2];
    __shared__ float4 sa[S1 * S1];
    __shared__ float2 sz[S1 * S1];
    __shared__ double4 lds[NT0 / 64];
    __shared__ uint32_t s_on[TT * TT];
    const int t = tid0();
    const int2 tq = tile_xy();
    const int x0 = tq.x * T0, y0 = tq.y * T0;
    if (t < TT * TT) s_on[t] = tile_on(c, F, (x0 >> 4) + t % TT, (y0 >> 4) + t / TT) ? 1u : 0u;
    const sp_cg_state* E = &c.K->st[par];
    if (st_done(*E)) return;
    __syncthreads();
    if (!any_on(s_on)) return;
    const float ax = (float)E->alpha[0], ay = (float)E->alpha[1];
    for (int k = t; k < S2 * S2; k += NT0) {
        const int gx = x0 - 2 + k % S2, gy = y0 - 2 + k / S2;
        float d = 0.0f;
        float2 rv = make_float2(0.0f, 0.0f);
        if (inside(F, gx, gy)) {
            const int i = at(F, gx, gy);
            d = c.mgd[i];
            rv = r_in[i];
            if (d > 0.0f) {
                const float2 q = c.cgq[i];
                rv.x -= ax * q.x;
                rv.y -= ay * q.y;
            }
        }
        sd[k] = d;
        sr[k] = rv;
    }
    for (int k = t; k < S1 * S1; k += NT0) {
        const int gx = x0 - 1 + k % S1, gy = y0 - 1 + k / S1;
        sa[k] = inside(F, gx, gy) ? c.mga[at(F, gx, gy)] : make_float4(0, 0, 0, 0);
    }
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    const bool mine = inside(F, x, y) && s_on[(threadIdx.y >> 4) * TT + (threadIdx.x >> 4)] != 0u;
    if (mine) {
        const int i = at(F, x, y);
        if (c.mgd[i] > 0.0f) {
            const float2 p = p_new[i];
            float2 x1 = c.pres[i];
            x1.x += ax * p.x;
            x1.y += ay * p.y;
            c.pres[i] = x1;
        }
    }
    __syncthreads();
    if (mine) r_out[at(F, x, y)] = sr[sx_(x, y, x0, y0, 2, S2)];
    down_z<NT0>(x0, y0, S1, S2, sd, sr, sa, sz);
    __syncthreads();
    if (mine) MG<H>::set_X(c, at(F, x, y), sz[sx_(x, y, x0, y0, 1, S1)]);
    if (!mine || ((threadIdx.x | threadIdx.y) & 1) != 0) return;
    const int cx = x >> 1, cy = y >> 1;
    if (!inside(C, cx, cy)) return;
    MG<H>::set_R(c, at(C, cx, cy), restrict_red(cx, cy, x0 - 1, y0 - 1, S1, S2, sd, sa, sz));
}

// mg_up.comp:38-75 on level 0: the coarse-grid correction and the black-then-red
// post-smoothing; with `reduce` also the partial sums r.z, r.r. r_in: the level-0 residual
// of this iteration (mgr, or r2 with rsel).
template <bool H>
__global__ __launch_bounds__(NT0) void k_up0(const TankBatch tb, Lv F, Lv C, int nbp, int gated, int par, int reduce, int rsel) {
    using M = MG<H>;
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const float2* __restrict__ r_in = rsel ? c.r2 : c.mgr;
    constexpr int S1 = T0 + 2, S2 = T0 + 4, SP = T0 / 2 + 2;
    __shared__ float sd[S2 * S2];
    __shared__ float2 sx[S2 * S2];
    __shared__ float4 sa[S1 * S1];
    __shared__ float2 sr[S1 * S1];
    __shared__ float2 szc[SP * SP];
    __shared__ float2 sb[S1 * S1];
    __shared__ double4 lds[NT0 / 64];
    __shared__ uint32_t s_on[TT * TT];
    const int t = tid0();
    const int2 tq = tile_xy();
    const int x0 = tq.x * T0, y0 = tq.y * T0;
    if (t < TT * TT) s_on[t] = tile_on(c, F, (x0 >> 4) + t % TT, (y0 >> 4) + t / TT) ? 1u : 0u;
    if (gated && slot_done(c, par)) return;
    __syncthreads();
    if (!any_on(s_on)) {
        if (reduce && t == 0) store_part(&c.part_sums[PART_RZ * nbp + slot0()], make_double4(0, 0, 0, 0));
        return;
    }
    for (int k = t; k < S2 * S2; k += NT0) {
        const int gx = x0 - 2 + k % S2, gy = y0 - 2 + k / S2;
        const bool in = inside(F, gx, gy);
        sd[k] = in ? M::D(c, 0, at(F, gx, gy)) : 0.0f;
        sx[k] = in ? M::X(c, at(F, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    for (int k = t; k < S1 * S1; k += NT0) {
        const int gx = x0 - 1 + k % S1, gy = y0 - 1 + k / S1;
        const bool in = inside(F, gx, gy);
        sa[k] = in ? M::A(c, 0, at(F, gx, gy)) : make_float4(0, 0, 0, 0);
        sr[k] = in ? r_in[at(F, gx, gy)] : make_float2(0.0f, 0.0f);
    }
    // parents of the ring-2 region: coarse cells x0/2 - 1 .. x0/2 + T0/2
    const int cx0 = (x0 >> 1) - 1, cy0 = (y0 >> 1) - 1;
    for (int k = t; k < SP * SP; k += NT0) {
        const int gx = cx0 + k % SP, gy = cy0 + k / SP;
        szc[k] = inside(C, gx, gy) ? M::Z(c, at(C, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    up_black<NT0>(x0, y0, 1, 2, S2, S1, cx0, cy0, SP, sd, sx, sa, sr, szc, sb);
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    double4 acc = make_double4(0, 0, 0, 0);
    if (inside(F, x, y) && s_on[(threadIdx.y >> 4) * TT + (threadIdx.x >> 4)]) {
        float2 rv;
        const float2 z = up_red(x, y, x0, y0, 1, S1, 2, S2, sd, sa, sr, sb, &rv);
        if (sd[sx_(x, y, x0, y0, 2, S2)] > 0.0f)
            acc = make_double4((double)rv.x * (double)z.x, (double)rv.y * (double)z.y, (double)rv.x * (double)rv.x,
                               (double)rv.y * (double)rv.y);
        M::set_Z(c, at(F, x, y), z);
    }
    if (reduce) {
        const double4 s2 = group_sum4<NT0>(acc, lds);
        if (t == 0) store_part(&c.part_sums[PART_RZ * nbp + slot0()], s2);
    }
}

// ---- coarse levels, one per launch (16 x 16 tiles) --------------------------------------

// mg_down.comp:36-69 on level l (F; its coarse level C): pre-smoothing from zero and the
// restriction of the residual. Level 0 (the first V-cycle of a solve) reads the CG residual
// (mgr, or r2 with rsel).
template <bool H>
__global__ __launch_bounds__(256) void k_mg_down(const TankBatch tb, Lv F, Lv C, int l, int gated, int par, int rsel) {
    using M = MG<H>;
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const float2* __restrict__ r_in = rsel ? c.r2 : c.mgr;
    constexpr int S1 = 18, S2 = 20;
    __shared__ float sd[S2 * S2];
    __shared__ float2 sr[S2 * S2];
    __shared__ float4 sa[S1 * S1];
    __shared__ float2 sz[S1 * S1];
    if (gated && slot_done(c, par)) return;
    const int2 tq = tile_xy();
    if (c.ptile[F.toff + slot_of(tq)] == 0u) return;
    const int x0 = tq.x * 16, y0 = tq.y * 16;
    const int t = tid2();
    for (int k = t; k < S2 * S2; k += 256) {
        const int gx = x0 - 2 + k % S2, gy = y0 - 2 + k / S2;
        const bool in = inside(F, gx, gy);
        sd[k] = in ? M::D(c, l, at(F, gx, gy)) : 0.0f;
        sr[k] = in ? (l == 0 ? r_in[at(F, gx, gy)] : M::R(c, at(F, gx, gy))) : make_float2(0.0f, 0.0f);
    }
    for (int k = t; k < S1 * S1; k += 256) {
        const int gx = x0 - 1 + k % S1, gy = y0 - 1 + k / S1;
        sa[k] = inside(F, gx, gy) ? M::A(c, l, at(F, gx, gy)) : make_float4(0, 0, 0, 0);
    }
    __syncthreads();
    down_z<256>(x0, y0, S1, S2, sd, sr, sa, sz);
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (inside(F, x, y)) M::set_X(c, at(F, x, y), sz[sx_(x, y, x0, y0, 1, S1)]);
    if (((threadIdx.x | threadIdx.y) & 1) != 0) return;
    const int cx = x >> 1, cy = y >> 1;
    if (!inside(C, cx, cy)) return;
    M::set_R(c, at(C, cx, cy), restrict_red(cx, cy, x0 - 1, y0 - 1, S1, S2, sd, sa, sz));
}

// mg_up.comp:38-75 on coarse level l >= 1 (F; its coarse level C).
template <bool H>
__global__ __launch_bounds__(256) void k_mg_up(const TankBatch tb, Lv F, Lv C, int l, int gated, int par) {
    using M = MG<H>;
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    constexpr int S1 = 18, S2 = 20, SP = 10;
    __shared__ float sd[S2 * S2];
    __shared__ float2 sx[S2 * S2];
    __shared__ float4 sa[S1 * S1];
    __shared__ float2 sr[S1 * S1];
    __shared__ float2 szc[SP * SP];
    __shared__ float2 sb[S1 * S1];
    if (gated && slot_done(c, par)) return;
    const int2 tq = tile_xy();
    if (c.ptile[F.toff + slot_of(tq)] == 0u) return;
    const int t = tid2();
    const int x0 = tq.x * 16, y0 = tq.y * 16;
    for (int k = t; k < S2 * S2; k += 256) {
        const int gx = x0 - 2 + k % S2, gy = y0 - 2 + k / S2;
        const bool in = inside(F, gx, gy);
        sd[k] = in ? M::D(c, l, at(F, gx, gy)) : 0.0f;
        sx[k] = in ? M::X(c, at(F, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    for (int k = t; k < S1 * S1; k += 256) {
        const int gx = x0 - 1 + k % S1, gy = y0 - 1 + k / S1;
        const bool in = inside(F, gx, gy);
        sa[k] = in ? M::A(c, l, at(F, gx, gy)) : make_float4(0, 0, 0, 0);
        sr[k] = in ? M::R(c, at(F, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    const int cx0 = (x0 >> 1) - 1, cy0 = (y0 >> 1) - 1;
    for (int k = t; k < SP * SP; k += 256) {
        const int gx = cx0 + k % SP, gy = cy0 + k / SP;
        szc[k] = inside(C, gx, gy) ? M::Z(c, at(C, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    up_black<256>(x0, y0, 1, 2, S2, S1, cx0, cy0, SP, sd, sx, sa, sr, szc, sb);
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (inside(F, x, y)) {
        float2 rv;
        M::set_Z(c, at(F, x, y), up_red(x, y, x0, y0, 1, S1, 2, S2, sd, sa, sr, sb, &rv));
    }
}

// ---- coarse levels two per launch ------------------------------------------------------
//
// A pair of adjacent coarse levels a = l and b = l + 1 in one launch: each work-group owns a
// 16 x 16 tile of level b and the 32 x 32 cells of level a under it, and works out the
// rings of both levels its stencils reach itself (the cells its neighbours also work out),
// so that no level waits for another launch. Every cell's arithmetic, and its order, is that
// of one level per launch; where a level's values pass through storage between the two
// launches (the right-hand side of b on the way down, z of b on the way up), they are
// rounded here as storage rounds them. So results are bit-identical to one level per launch.

namespace {

constexpr int NTF = 1024;  // work-items of a two-level kernel (32 x 32)

// Way down: level a's diagonal and right-hand side on 6 rings, couplings and z on 5; level
// b's diagonal and right-hand side on 2, couplings and z on 1.
constexpr int D2_AD = 44, D2_AA = 42, D2_BD = 20, D2_BA = 18;
// Way up: level b's diagonal and iterate on 3 rings, couplings, right-hand side and black
// values on 2, z on 1, its parents on 12 x 12; level a's diagonal and iterate on 2,
// couplings, right-hand side and black values on 1.
constexpr int U2_BX = 22, U2_BR = 20, U2_BZ = 18, U2_BP = 12, U2_AX = 36, U2_AR = 34;

}  // namespace

constexpr size_t DOWN2_LDS = sizeof(float4) * (D2_AA * D2_AA + D2_BA * D2_BA) +
                             sizeof(float2) * (D2_AD * D2_AD + D2_AA * D2_AA + D2_BD * D2_BD + D2_BA * D2_BA) +
                             sizeof(float) * (D2_AD * D2_AD + D2_BD * D2_BD);
constexpr size_t UP2_LDS = sizeof(float4) * (U2_BR * U2_BR + U2_AR * U2_AR) +
                           sizeof(float2) * (U2_BX * U2_BX + 2 * U2_BR * U2_BR + U2_BP * U2_BP + U2_BZ * U2_BZ +
                                             U2_AX * U2_AX + 2 * U2_AR * U2_AR) +
                           sizeof(float) * (U2_BX * U2_BX + U2_AX * U2_AX);

// mg_down.comp on levels a = l and b = l + 1 (A, B; B's coarse level Cc).
template <bool H>
__global__ __launch_bounds__(NTF) void k_mg_down2(const TankBatch tb, Lv A, Lv B, Lv Cc, int la, int gated, int par) {
    using M = MG<H>;
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    extern __shared__ float4 dyn_lds[];
    float4* aA = dyn_lds;
    float4* bA = aA + D2_AA * D2_AA;
    float2* ar = reinterpret_cast<float2*>(bA + D2_BA * D2_BA);
    float2* az = ar + D2_AD * D2_AD;
    float2* bR = az + D2_AA * D2_AA;
    float2* bz = bR + D2_BD * D2_BD;
    float* aD = reinterpret_cast<float*>(bz + D2_BA * D2_BA);
    float* bD = aD + D2_AD * D2_AD;
    __shared__ uint32_t s_on[4];  // level a's four tiles under this block
    const int lb = la + 1;
    if (gated && slot_done(c, par)) return;
    const int2 tq = tile_xy();
    if (c.ptile[B.toff + slot_of(tq)] == 0u) return;
    const int t = threadIdx.y * 32 + threadIdx.x;
    const int X0 = tq.x * 16, Y0 = tq.y * 16, ax0 = 2 * X0, ay0 = 2 * Y0;
    if (t < 4) s_on[t] = tile_on(c, A, (ax0 >> 4) + (t & 1), (ay0 >> 4) + (t >> 1)) ? 1u : 0u;
    for (int k = t; k < D2_AD * D2_AD; k += NTF) {
        const int gx = ax0 - 6 + k % D2_AD, gy = ay0 - 6 + k / D2_AD;
        const bool in = inside(A, gx, gy);
        aD[k] = in ? M::D(c, la, at(A, gx, gy)) : 0.0f;
        ar[k] = in ? M::R(c, at(A, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    for (int k = t; k < D2_AA * D2_AA; k += NTF) {
        const int gx = ax0 - 5 + k % D2_AA, gy = ay0 - 5 + k / D2_AA;
        aA[k] = inside(A, gx, gy) ? M::A(c, la, at(A, gx, gy)) : make_float4(0, 0, 0, 0);
    }
    for (int k = t; k < D2_BD * D2_BD; k += NTF) {
        const int gx = X0 - 2 + k % D2_BD, gy = Y0 - 2 + k / D2_BD;
        bD[k] = inside(B, gx, gy) ? M::D(c, lb, at(B, gx, gy)) : 0.0f;
    }
    for (int k = t; k < D2_BA * D2_BA; k += NTF) {
        const int gx = X0 - 1 + k % D2_BA, gy = Y0 - 1 + k / D2_BA;
        bA[k] = inside(B, gx, gy) ? M::A(c, lb, at(B, gx, gy)) : make_float4(0, 0, 0, 0);
    }
    __syncthreads();
    // level a: z on its region and five rings (the region of its four-ring arrays is ax0 - 5)
    down_z<NTF>(ax0 - 4, ay0 - 4, D2_AA, D2_AD, aD, ar, aA, az);
    __syncthreads();
    for (int k = t; k < 32 * 32; k += NTF) {
        const int lx = k & 31, ly = k >> 5, gx = ax0 + lx, gy = ay0 + ly;
        if (inside(A, gx, gy) && s_on[(ly >> 4) * 2 + (lx >> 4)]) M::set_X(c, at(A, gx, gy), az[sx_(gx, gy, ax0, ay0, 5, D2_AA)]);
    }
    // level b's right-hand side on two rings, as level b reads it back from storage
    for (int k = t; k < D2_BD * D2_BD; k += NTF) {
        const int qx = X0 - 2 + k % D2_BD, qy = Y0 - 2 + k / D2_BD;
        float2 sum = make_float2(0.0f, 0.0f);
        if (inside(B, qx, qy)) sum = restrict_red(qx, qy, ax0 - 5, ay0 - 5, D2_AA, D2_AD, aD, aA, az);
        bR[k] = M::stored_R(sum);
        const int ox = qx - X0, oy = qy - Y0;
        if (ox >= 0 && oy >= 0 && ox < 16 && oy < 16 && inside(B, qx, qy) && s_on[(oy >> 3) * 2 + (ox >> 3)])
            M::set_R(c, at(B, qx, qy), sum);
    }
    __syncthreads();
    // level b: z on its tile and one ring
    down_z<NTF>(X0, Y0, D2_BA, D2_BD, bD, bR, bA, bz);
    __syncthreads();
    for (int k = t; k < 16 * 16; k += NTF) {
        const int lx = k & 15, ly = k >> 4, x = X0 + lx, y = Y0 + ly;
        if (inside(B, x, y)) M::set_X(c, at(B, x, y), bz[sx_(x, y, X0, Y0, 1, D2_BA)]);
        if (((lx | ly) & 1) != 0) continue;
        const int cx = x >> 1, cy = y >> 1;
        if (!inside(Cc, cx, cy)) continue;
        M::set_R(c, at(Cc, cx, cy), restrict_red(cx, cy, X0 - 1, Y0 - 1, D2_BA, D2_BD, bD, bA, bz));
    }
}

// mg_up.comp on levels b = l + 1 and then a = l (A, B; B's coarse level Cc holds z already).
template <bool H>
__global__ __launch_bounds__(NTF) void k_mg_up2(const TankBatch tb, Lv A, Lv B, Lv Cc, int la, int gated, int par) {
    using M = MG<H>;
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    extern __shared__ float4 dyn_lds[];
    float4* bA = dyn_lds;
    float4* aA = bA + U2_BR * U2_BR;
    float2* bx = reinterpret_cast<float2*>(aA + U2_AR * U2_AR);
    float2* br = bx + U2_BX * U2_BX;
    float2* bsb = br + U2_BR * U2_BR;
    float2* bzc = bsb + U2_BR * U2_BR;
    float2* bz = bzc + U2_BP * U2_BP;
    float2* ax = bz + U2_BZ * U2_BZ;
    float2* ar = ax + U2_AX * U2_AX;
    float2* asb = ar + U2_AR * U2_AR;
    float* bD = reinterpret_cast<float*>(asb + U2_AR * U2_AR);
    float* aD = bD + U2_BX * U2_BX;
    __shared__ uint32_t s_on[4];
    const int lb = la + 1;
    if (gated && slot_done(c, par)) return;
    const int2 tq = tile_xy();
    if (c.ptile[B.toff + slot_of(tq)] == 0u) return;
    const int t = threadIdx.y * 32 + threadIdx.x;
    const int X0 = tq.x * 16, Y0 = tq.y * 16, ax0 = 2 * X0, ay0 = 2 * Y0;
    if (t < 4) s_on[t] = tile_on(c, A, (ax0 >> 4) + (t & 1), (ay0 >> 4) + (t >> 1)) ? 1u : 0u;
    for (int k = t; k < U2_BX * U2_BX; k += NTF) {
        const int gx = X0 - 3 + k % U2_BX, gy = Y0 - 3 + k / U2_BX;
        const bool in = inside(B, gx, gy);
        bD[k] = in ? M::D(c, lb, at(B, gx, gy)) : 0.0f;
        bx[k] = in ? M::X(c, at(B, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    for (int k = t; k < U2_BR * U2_BR; k += NTF) {
        const int gx = X0 - 2 + k % U2_BR, gy = Y0 - 2 + k / U2_BR;
        const bool in = inside(B, gx, gy);
        bA[k] = in ? M::A(c, lb, at(B, gx, gy)) : make_float4(0, 0, 0, 0);
        br[k] = in ? M::R(c, at(B, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    // parents of level b's three-ring region: coarse cells X0/2 - 2 .. X0/2 + 9
    const int pcx = (X0 >> 1) - 2, pcy = (Y0 >> 1) - 2;
    for (int k = t; k < U2_BP * U2_BP; k += NTF) {
        const int gx = pcx + k % U2_BP, gy = pcy + k / U2_BP;
        bzc[k] = inside(Cc, gx, gy) ? M::Z(c, at(Cc, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    for (int k = t; k < U2_AX * U2_AX; k += NTF) {
        const int gx = ax0 - 2 + k % U2_AX, gy = ay0 - 2 + k / U2_AX;
        const bool in = inside(A, gx, gy);
        aD[k] = in ? M::D(c, la, at(A, gx, gy)) : 0.0f;
        ax[k] = in ? M::X(c, at(A, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    for (int k = t; k < U2_AR * U2_AR; k += NTF) {
        const int gx = ax0 - 1 + k % U2_AR, gy = ay0 - 1 + k / U2_AR;
        const bool in = inside(A, gx, gy);
        aA[k] = in ? M::A(c, la, at(A, gx, gy)) : make_float4(0, 0, 0, 0);
        ar[k] = in ? M::R(c, at(A, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    // level b: black values on two rings, then z on one ring (its own tile written)
    up_black<NTF>(X0, Y0, 2, 3, U2_BX, U2_BR, pcx, pcy, U2_BP, bD, bx, bA, br, bzc, bsb);
    __syncthreads();
    for (int k = t; k < U2_BZ * U2_BZ; k += NTF) {
        const int gx = X0 - 1 + k % U2_BZ, gy = Y0 - 1 + k / U2_BZ;
        float2 rv;
        const float2 z = up_red(gx, gy, X0, Y0, 2, U2_BR, 3, U2_BX, bD, bA, br, bsb, &rv);
        const int ox = gx - X0, oy = gy - Y0;
        if (ox >= 0 && oy >= 0 && ox < 16 && oy < 16 && inside(B, gx, gy)) M::set_Z(c, at(B, gx, gy), z);
        bz[k] = M::stored_Z(z);  // as level a reads it back from storage
    }
    __syncthreads();
    // level a: black values on one ring from the parents' z, then z on its region
    up_black<NTF>(ax0, ay0, 1, 2, U2_AX, U2_AR, X0 - 1, Y0 - 1, U2_BZ, aD, ax, aA, ar, bz, asb);
    __syncthreads();
    for (int k = t; k < 32 * 32; k += NTF) {
        const int lx = k & 31, ly = k >> 5, gx = ax0 + lx, gy = ay0 + ly;
        if (!inside(A, gx, gy) || !s_on[(ly >> 4) * 2 + (lx >> 4)]) continue;
        float2 rv;
        M::set_Z(c, at(A, gx, gy), up_red(gx, gy, ax0, ay0, 1, U2_AR, 2, U2_AX, aD, aA, ar, asb, &rv));
    }
}

// mg_tail.comp mode 0 (mg_tail.comp:159-225): every level from the tail down in one
// 1024 work-item group, with the iterate, diagonal and coarse right-hand sides in LDS.
// Levels of more than 128 cells are swept by the whole group, smaller ones by its first
// wave64 alone (the arithmetic and its order per cell are the same either way).
#define WSYNC()                                                  \
    do {                                                         \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   \
        __builtin_amdgcn_wave_barrier();                         \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");   \
    } while (0)

#ifndef SP_TAIL_NT
#define SP_TAIL_NT 1024
#endif
constexpr int TAIL_NT = SP_TAIL_NT;  // work-items of the tail's group

// LDS (dynamic, sized at launch to the hierarchy): per tail cell its couplings (float4),
// iterate and right-hand side (float2) and diagonal; the first level's right-hand side is
// loaded once with the matrix, so that the sweeps never wait on global memory.
template <bool H>
__global__ __launch_bounds__(TAIL_NT) void k_mg_tail(const TankBatch tb, int gated, int par) {
    using M = MG<H>;
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    extern __shared__ float4 tail_lds[];
    __shared__ int4 sl[SP_MAX_LEVELS];
    if (gated && slot_done(c, par)) return;
    const int first = c.tail, last = c.n_levels - 1;
    const int4 Lf = c.lev[first], Le = c.lev[last];
    const int BASE = Lf.x;
    const int total = Le.x + Le.y * Le.z - BASE;
    const int NF = Lf.y * Lf.z;
    float4* sa = tail_lds;
    float2* sx = reinterpret_cast<float2*>(sa + total);
    float2* sr = sx + total;
    float* sd = reinterpret_cast<float*>(sr + total);
    const int bt = threadIdx.x;
    if (bt <= last) sl[bt] = c.lev[bt];
    __syncthreads();
    for (int t = bt; t < total; t += TAIL_NT) {
        int l = first;  // the level of tail cell t (the storage of the coarse matrices is per level)
        if (H)
            while (l < last && BASE + t >= sl[l + 1].x) l++;
        sd[t] = M::D(c, l, BASE + t);
        sa[t] = M::A(c, l, BASE + t);
        if (t < NF) sr[t] = M::R(c, BASE + t);
    }
    __syncthreads();

    // The neighbours' operands are loaded without waiting on whether each is an unknown, and
    // added in order where it is: the same sums as adding under a branch, without a chain of
    // dependent LDS round trips per cell.
    auto sat = [&](int4 L, int x, int y) { return L.x - BASE + y * L.y + x; };
    // slot of cell (x, y) of level L, clamped to the level's first cell outside it; whether it
    // is an unknown
    auto nbr = [&](int4 L, int x, int y, int& s) -> bool {
        const bool in = x >= 0 && y >= 0 && x < L.y && y < L.z;
        s = L.x - BASE + (in ? y * L.y + x : 0);
        const float dn = sd[s];
        return in & (dn > 0.0f);
    };
    auto sweep = [&](int4 L, int colour, bool zero, int t0, int step) {
        for (int t = t0; t < L.y * L.z; t += step) {
            const int x = t % L.y, y = t / L.y;
            const int s = L.x - BASE + t;
            const float D = sd[s];
            if (D <= 0.0f || ((x + y) & 1) != colour) {
                if (zero) sx[s] = make_float2(0.0f, 0.0f);
                continue;
            }
            float2 sum = sr[s];
            if (!zero) {
                const float4 w = sa[s];
                int sn[4];
                bool ok[4];
                float2 xn[4];
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    ok[d] = nbr(L, x + NB_X[d], y + NB_Y[d], sn[d]);
                    xn[d] = sx[sn[d]];
                }
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    const float wd = f4c(w, d);
                    const float2 cand = make_float2(sum.x + wd * xn[d].x, sum.y + wd * xn[d].y);
                    sum = ok[d] ? cand : sum;
                }
            }
            sx[s] = make_float2(sum.x / D, sum.y / D);
        }
    };
    auto restrict_to = [&](int4 F, int4 C, int t0, int step) {
        for (int t = t0; t < C.y * C.z; t += step) {
            const int cx = t % C.y, cy = t / C.y;
            float2 sum = make_float2(0.0f, 0.0f);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int x = 2 * cx + (k & 1), y = 2 * cy + (k >> 1);
                int s;
                const bool okc = nbr(F, x, y, s);
                const float4 w = sa[s];
                const float Dc = sd[s];
                const float2 xc = sx[s];
                const float2 r = sr[s];
                int sn[4];
                bool ok[4];
                float2 xn[4];
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    ok[d] = nbr(F, x + NB_X[d], y + NB_Y[d], sn[d]);
                    xn[d] = sx[sn[d]];
                }
                float2 az = make_float2(Dc * xc.x, Dc * xc.y);
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    const float wd = f4c(w, d);
                    const float2 cand = make_float2(az.x - wd * xn[d].x, az.y - wd * xn[d].y);
                    az = ok[d] ? cand : az;
                }
                const float2 cand = make_float2(sum.x + (r.x - az.x), sum.y + (r.y - az.y));
                sum = okc ? cand : sum;
            }
            sr[C.x - BASE + t] = sum;
        }
    };
    auto post_black = [&](int4 F, int4 C, int t0, int step) {
        for (int t = t0; t < F.y * F.z; t += step) {
            const int x = t % F.y, y = t / F.y;
            if (((x + y) & 1) == 0) continue;
            const int s = F.x - BASE + t;
            const float D = sd[s];
            if (D <= 0.0f) continue;
            const float4 w = sa[s];
            float2 sum = sr[s];
            bool ok[4];
            float2 xn[4], zc[4];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int nx = x + NB_X[d], ny = y + NB_Y[d];
                int sn;
                ok[d] = nbr(F, nx, ny, sn);
                xn[d] = sx[sn];
                zc[d] = sx[ok[d] ? sat(C, nx >> 1, ny >> 1) : C.x - BASE];
            }
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const float wd = f4c(w, d);
                const float2 cand = make_float2(sum.x + wd * (xn[d].x + SP_MG_OMEGA * zc[d].x),
                                                sum.y + wd * (xn[d].y + SP_MG_OMEGA * zc[d].y));
                sum = ok[d] ? cand : sum;
            }
            sx[s] = make_float2(sum.x / D, sum.y / D);
        }
    };

    int ls = first;
    while (ls < last && sl[ls].y * sl[ls].z > SP_TAIL_WARP_CELLS) ls++;
    for (int l = first; l < ls; l++) {
        sweep(sl[l], 0, true, bt, TAIL_NT);
        __syncthreads();
        sweep(sl[l], 1, false, bt, TAIL_NT);
        __syncthreads();
        restrict_to(sl[l], sl[l + 1], bt, TAIL_NT);
        __syncthreads();
    }
    if (bt < 64) {
        const int wt = bt, ws = 64;
        for (int l = ls; l < last; l++) {
            sweep(sl[l], 0, true, wt, ws);
            WSYNC();
            sweep(sl[l], 1, false, wt, ws);
            WSYNC();
            restrict_to(sl[l], sl[l + 1], wt, ws);
            WSYNC();
        }
        sweep(sl[last], 0, true, wt, ws);
        WSYNC();
        for (int h = 1; h < SP_TAIL_HALF_SWEEPS; h++) {
            sweep(sl[last], h & 1, false, wt, ws);
            WSYNC();
        }
        for (int l = last - 1; l >= ls; l--) {
            post_black(sl[l], sl[l + 1], wt, ws);
            WSYNC();
            sweep(sl[l], 0, false, wt, ws);
            WSYNC();
        }
    }
    __syncthreads();
    for (int l = ls - 1; l >= first; l--) {
        post_black(sl[l], sl[l + 1], bt, TAIL_NT);
        __syncthreads();
        sweep(sl[l], 0, false, bt, TAIL_NT);
        __syncthreads();
    }
    for (int t = bt; t < NF; t += TAIL_NT) M::set_Z(c, BASE + t, sx[t]);
}


// The end of a solve (reduce_tail.glsl:105-116, modes 3 and 9): the scalar step of its last
// iteration (or its start, when none ran), the final state into the record's first fields,
// and the statistics; the second solve keeps its own iteration count.
__global__ __launch_bounds__(REDUCE_NT) void k_cg_end(const TankBatch tb, int nbp, int iters, int second) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ double4 lds[REDUCE_NT / 64];
    double4 a = make_double4(0, 0, 0, 0), b = a;
    const sp_cg_state* P = &c.K->st[(iters - 1) & 1];
    const bool ended = iters > 0 && st_done(*P);
    if (!ended) a = total_of<REDUCE_NT>(c.part_sums, PART_RZ * nbp, nbp, lds);
    if (iters == 0) b = total_of<REDUCE_NT>(c.part_sums, PART_BB * nbp, nbp, lds);
    if (threadIdx.x != 0) return;
    sp_cg_state E;
    if (iters == 0) {
        E = cg_start(c, a, b, second != 0);
        c.K->solves += 1u;
    } else {
        E = ended ? *P : cg_next(c, *P, a);
    }
    sp_solve* K = c.K;
    for (int i = 0; i < 2; i++) {
        K->rz[i] = E.rz[i];
        K->rr[i] = E.rr[i];
        K->bb[i] = E.bb[i];
        K->beta[i] = E.beta[i];
        K->done[i] = E.done[i];
        K->iters[i] = E.iters[i];
    }
    K->sweeps = E.sweeps;
    K->halt = E.halt;
    K->cap_now = E.cap;
    const uint32_t it = E.iters[0] > E.iters[1] ? E.iters[0] : E.iters[1];
    if (second) {
        if (E.iters[0] > K->p2_it_max) K->p2_it_max = E.iters[0];
        if (E.done[0] == 0u) K->p2_unconv += 1u;
        K->p2_iters_last = E.iters[0];
    } else {
        if (it > K->it_max) K->it_max = it;
        if (E.done[0] == 0u) K->unconv += 1u;
        K->iters_last = it;
        K->iters_sum += it;
    }
    const float rel = E.bb[0] > 0.0 ? (float)sqrt(E.rr[0] / E.bb[0]) : 0.0f;
    if (rel > K->res_max) K->res_max = rel;
}
