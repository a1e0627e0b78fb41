// solve.hip — the pressure solve on gfx950: conjugate gradients on two right-hand sides
// at once (pressure * dt and the density-correction potential), preconditioned by an
// aggregation multigrid V-cycle (mod.rs:2929-3010; mg_*.comp, cg_*.comp).
//
// Layout: every level of the hierarchy back to back in mga (float4 couplings L, R, B, T),
// mgd (diagonal), mgx / mgz / mgr (float2 per cell: pre-smoothed iterate, finished
// iterate, right-hand side), as the reference lays them out (mg.glsl:1-37). Level kernels
// run one work-item per cell on 16 x 16 tiles and skip tiles without unknowns (ptile).
//
// Precision: cell arithmetic is FP32 in the reference's operation order; the scalar
// reductions (r.z, r.r, p.q, b.b) are FP64 — the product of two floats is exact in double
// and the sums are formed per work-group and then in a fixed order by a one-group kernel,
// so the scalars do not depend on scheduling. With mg_half the V-cycle's storage is
// narrower (MG<true> below); its arithmetic stays FP32.
//
// Convergence without host round trips: the scalar state lives in device memory (sp_solve);
// every kernel of an iteration reads the two `done` flags at entry and returns at once when
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

// Storage of the multigrid levels. MG<false>: FP32, the reference's layout. MG<true>
// (mg_half): every coarse level (1 and below) narrow: iterates (mgx, mgz) and right-hand
// sides (mgr) in BF16, two per 32-bit word; couplings and diagonals in FP16, scaled by 2^-l
// so that their growth with the level (about 2x per level under 2 x 2 aggregation) stays
// far from FP16's range. Level 0 stays FP32: its matrix is also the CG operator, its
// residual a CG vector, and its finished iterate z the preconditioned residual CG steps
// along (in BF16 that last one cost 1.8x the iterations at 1024 x 576). Narrow values sit at
// the start of each array's coarse region, which begins at cell N (level 1). Loads widen to
// FP32 exactly, so all arithmetic is FP32; stores round to nearest even.
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
    SP_DEV static float2 X(const Ctx& c, int i) { return vec(c.mgx, c, i); }
    SP_DEV static void set_X(const Ctx& c, int i, float2 v) { set_vec(c.mgx, c, i, v); }
    SP_DEV static float2 Z(const Ctx& c, int i) { return vec(c.mgz, c, i); }
    SP_DEV static void set_Z(const Ctx& c, int i, float2 v) { set_vec(c.mgz, c, i, v); }
    SP_DEV static float2 R(const Ctx& c, int i) {
        return unpack_bf2(reinterpret_cast<const uint32_t*>(c.mgr + c.N)[i - c.N]);
    }
    SP_DEV static void set_R(const Ctx& c, int i, float2 v) { reinterpret_cast<uint32_t*>(c.mgr + c.N)[i - c.N] = pack_bf2(v); }
};
// A solve is over for this tank when both potentials converged or when it has done the
// iterations this tank records (its cap), whichever comes first.
SP_DEV bool solve_done(const Ctx& c) { return (c.K->done[0] != 0u && c.K->done[1] != 0u) || c.K->halt != 0u; }

SP_DEV int tid2() { return threadIdx.y * 16 + threadIdx.x; }
SP_DEV int tile_slot() { return blockIdx.y * gridDim.x + blockIdx.x; }

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
#pragma unroll
    for (int w = 1; w < NT / 64; w++) add4(s, lds[w]);
    __syncthreads();
    return s;
}

SP_DEV void store_part(double4* p, double4 v) { *p = v; }

// Sum of parts[base .. base + n) by one group of NT work-items in a fixed order; each
// work-item issues its loads four at a time before adding, so their latencies overlap.
template <int NT>
SP_DEV double4 total_of(const double4* parts, int base, int n, double4* lds) {
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    double4 s = make_double4(0, 0, 0, 0);
    for (int k0 = 0; k0 < n; k0 += 4 * NT) {
        double4 v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + j * NT + t;
            v[j] = k < n ? parts[base + k] : make_double4(0, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) add4(s, v[j]);
    }
    return group_sum4<NT>(s, lds);
}
constexpr int REDUCE_NT = 256;  // work-items of the one-group reduction kernels
SP_DEV double4 total(const Ctx& c, int base, int n, double4* lds) { return total_of<REDUCE_NT>(c.part_sums, base, n, lds); }

// Modes 0, 1, 2 of reduce_tail.glsl, and 8: the start of the second solve (after the
// viscosity), where the density-correction potential counts as solved.
enum { RED_START = 0, RED_ALPHA = 1, RED_BETA = 2, RED_START2 = 8 };

// The scalar step of the pressure CG (reduce_tail.glsl:59-116 modes 0, 1, 2), FP64.
// parts [0, n): r.z (xy), r.r (zw) of the last V-cycle; [n, 2n): b.b (xy) or p.q (xy).
SP_DEV void reduce_scalars(const Ctx& c, int mode, int n, double4* lds) {
    sp_solve* K = c.K;
    const sp_gparams& P = *c.P;
    const double tx = (double)P.tol[0] * (double)P.tol[0], ty = (double)P.tol[1] * (double)P.tol[1];
    if (mode == RED_START || mode == RED_START2) {
        double4 a = total(c, 0, n, lds);
        double4 b = total(c, n, n, lds);
        if (tid2() == 0) {
            K->bb[0] = b.x;
            K->bb[1] = b.y;
            K->rz[0] = a.x;
            K->rz[1] = a.y;
            K->rr[0] = a.z;
            K->rr[1] = a.w;
            K->beta[0] = K->beta[1] = 0.0;
            K->alpha[0] = K->alpha[1] = 0.0;
            K->iters[0] = K->iters[1] = 0u;
            K->sweeps = 0u;
            K->halt = 0u;
            K->cap_now = mode == RED_START2 ? K->cap_p2 : K->cap_p;
            K->done[0] = a.z <= tx * b.x ? 1u : 0u;
            K->done[1] = (mode == RED_START2 || a.w <= ty * b.y) ? 1u : 0u;
            K->solves += 1u;
        }
    } else if (mode == RED_ALPHA) {
        double4 b = total(c, n, n, lds);
        if (tid2() == 0) {
            K->alpha[0] = (K->done[0] == 0u && b.x > 0.0) ? K->rz[0] / b.x : 0.0;
            K->alpha[1] = (K->done[1] == 0u && b.y > 0.0) ? K->rz[1] / b.y : 0.0;
        }
    } else {
        double4 a = total(c, 0, n, lds);
        if (tid2() == 0) {
            uint32_t was0 = K->done[0], was1 = K->done[1];
            K->iters[0] += 1u - was0;
            K->iters[1] += 1u - was1;
            uint32_t d0 = was0 | (a.z <= tx * K->bb[0] ? 1u : 0u);
            uint32_t d1 = was1 | (a.w <= ty * K->bb[1] ? 1u : 0u);
            K->beta[0] = (d0 == 0u && K->rz[0] > 0.0) ? a.x / K->rz[0] : 0.0;
            K->beta[1] = (d1 == 0u && K->rz[1] > 0.0) ? a.y / K->rz[1] : 0.0;
            K->rz[0] = a.x;
            K->rz[1] = a.y;
            if (!was0) K->rr[0] = a.z;
            if (!was1) K->rr[1] = a.w;
            K->done[0] = d0;
            K->done[1] = d1;
            // the last kernel of an iteration: stop here if this tank records no more
            K->sweeps += 1u;
            if (K->cap_now != 0u && K->sweeps >= K->cap_now && (d0 == 0u || d1 == 0u)) K->halt = 1u;
        }
    }
}

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

// cg_init.comp: r = b - A x and the partial sums of b.b (into the second half of parts).
__global__ __launch_bounds__(256) void k_cg_init(const TankBatch tb, int nblocks) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ double4 lds[4];
    const int slot = tile_slot();
    if (c.ptile[slot] == 0u) {
        if (tid2() == 0) store_part(&c.part_sums[nblocks + slot], make_double4(0, 0, 0, 0));
        return;
    }
    const Lv F = level(c, 0);
    int x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    double4 acc = make_double4(0, 0, 0, 0);
    if (inside(F, x, y)) {
        int i = at(F, x, y);
        float D = c.mgd[i];
        if (D > 0.0f) {
            float2 xv = c.pres[i];
            float4 w = c.mga[i];
            float2 ax = make_float2(D * xv.x, D * xv.y);
            for (int d = 0; d < 4; d++) {
                int nx = x + NB_X[d], ny = y + NB_Y[d];
                if (unknown(c, F, nx, ny)) {
                    float wd = f4c(w, d);
                    float2 pn = c.pres[at(F, nx, ny)];
                    ax.x -= wd * pn.x;
                    ax.y -= wd * pn.y;
                }
            }
            float4 b = c.prhs[i];
            c.mgr[i] = make_float2(b.x - ax.x, b.y - ax.y);
            acc.x = (double)b.x * (double)b.x;
            acc.y = (double)b.y * (double)b.y;
        } else {
            c.mgr[i] = make_float2(0.0f, 0.0f);
        }
    }
    double4 s = group_sum4(acc, lds);
    if (tid2() == 0) store_part(&c.part_sums[nblocks + slot], s);
}

// ---- LDS-staged level kernels ----------------------------------------------------------
//
// Each work-group owns a 16 x 16 tile of its level. It first loads what its stencils reach
// (the tile plus a ring of one or two cells: diagonals, couplings, right-hand sides,
// iterates) into LDS with every work-item issuing its loads at once, then computes from
// LDS. The arithmetic per cell, and its order, is that of the reference's kernels (which
// read neighbours from global memory and recompute values they do not hold); only where
// the operands come from changes, so results are bit-identical.

namespace {

constexpr int R1 = 18;  // tile + 1-cell ring
constexpr int R2 = 20;  // tile + 2-cell ring

// Index of global cell (x, y) in a ring-r staging array of the tile at (x0, y0).
SP_DEV int sidx(int x, int y, int x0, int y0, int r) { return (y - y0 + r) * (16 + 2 * r) + (x - x0 + r); }

SP_DEV float2 direction_of(float2 z, float2 po, float bx, float by) {
    if (bx == 0.0f && by == 0.0f) return z;
    return make_float2(bx != 0.0f ? z.x + bx * po.x : z.x, by != 0.0f ? z.y + by * po.y : z.y);
}

}  // namespace

// cg_apply.comp: p = z + beta p_old, q = A p, and the partial sums of p.q (second half of
// the partial buffer).
template <bool H>
__global__ __launch_bounds__(256) void k_cg_apply(const TankBatch tb, int nblocks, int par) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    // the search direction alternates between cgp and cgp2 with the iteration's parity
    const float2* __restrict__ p_old = par ? c.cgp2 : c.cgp;
    float2* __restrict__ p_new = par ? c.cgp : c.cgp2;
    __shared__ double4 lds[4];
    __shared__ float sd[R1 * R1];
    __shared__ float2 sp[R1 * R1];
    if (solve_done(c)) return;
    const int slot = tile_slot();
    if (c.ptile[slot] == 0u) {
        if (tid2() == 0) store_part(&c.part_sums[nblocks + slot], make_double4(0, 0, 0, 0));
        return;
    }
    const Lv F = level(c, 0);
    const float bx = (float)c.K->beta[0], by = (float)c.K->beta[1];
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    const int t = tid2();
    for (int k = t; k < R1 * R1; k += 256) {
        int gx = x0 - 1 + k % R1, gy = y0 - 1 + k / R1;
        float d = 0.0f;
        float2 pv = make_float2(0.0f, 0.0f);
        if (inside(F, gx, gy)) {
            int i = at(F, gx, gy);
            d = c.mgd[i];
            if (d > 0.0f) {
                float2 po = (bx == 0.0f && by == 0.0f) ? make_float2(0.0f, 0.0f) : p_old[i];
                pv = direction_of(MG<H>::Z(c, i), po, bx, by);
            }
        }
        sd[k] = d;
        sp[k] = pv;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    double4 acc = make_double4(0, 0, 0, 0);
    if (inside(F, x, y)) {
        int i = at(F, x, y);
        int si = sidx(x, y, x0, y0, 1);
        float D = sd[si];
        if (D > 0.0f) {
            float2 p = sp[si];
            float4 w = c.mga[i];
            float2 q = make_float2(D * p.x, D * p.y);
            for (int d = 0; d < 4; d++) {
                int sn = sidx(x + NB_X[d], y + NB_Y[d], x0, y0, 1);
                if (sd[sn] > 0.0f) {
                    float wd = f4c(w, d);
                    float2 pn = sp[sn];
                    q.x -= wd * pn.x;
                    q.y -= wd * pn.y;
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
    double4 s = group_sum4(acc, lds);
    if (t == 0) store_part(&c.part_sums[nblocks + slot], s);
}

namespace {

// The way down on one level (mg_down.comp:36-69) over a tile whose residual for the tile
// plus two rings is in sr (ring 2) and diagonals in sd (ring 2): z of the tile and its first
// ring, then the restriction of the red cells' residual to the coarse level.
struct DownStage {
    float sd[R2 * R2];
    float2 sr[R2 * R2];
    float4 sa[R1 * R1];
    float2 sz[R1 * R1];
};

template <bool H>
SP_DEV void down_compute(const Ctx& c, const Lv& F, const Lv& C, int x0, int y0, DownStage& S) {
    const int t = tid2();
    // z after the red half-sweep from zero: red z = r / D; black from its red neighbours.
    for (int k = t; k < R1 * R1; k += 256) {
        int gx = x0 - 1 + k % R1, gy = y0 - 1 + k / R1;
        int s2 = sidx(gx, gy, x0, y0, 2);
        float D = S.sd[s2];
        float2 z = make_float2(0.0f, 0.0f);
        if (D > 0.0f) {
            float2 r = S.sr[s2];
            if (((gx + gy) & 1) == 0) {
                z = make_float2(r.x / D, r.y / D);
            } else {
                float4 w = S.sa[k];
                float2 sum = r;
                for (int d = 0; d < 4; d++) {
                    int sn = sidx(gx + NB_X[d], gy + NB_Y[d], x0, y0, 2);
                    float dj = S.sd[sn];
                    if (!(dj > 0.0f)) continue;
                    float wd = f4c(w, d);
                    float2 rj = S.sr[sn];
                    sum.x += wd * rj.x / dj;
                    sum.y += wd * rj.y / dj;
                }
                z = make_float2(sum.x / D, sum.y / D);
            }
        }
        S.sz[k] = z;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (inside(F, x, y)) MG<H>::set_X(c, at(F, x, y), S.sz[sidx(x, y, x0, y0, 1)]);
    if (((threadIdx.x | threadIdx.y) & 1) != 0) return;
    const int cx = x >> 1, cy = y >> 1;
    if (!inside(C, cx, cy)) return;
    float2 sum = make_float2(0.0f, 0.0f);
    for (int k = 0; k < 2; k++) {
        int qx = x + k, qy = y + k;
        int sq = sidx(qx, qy, x0, y0, 1);
        if (!(S.sd[sidx(qx, qy, x0, y0, 2)] > 0.0f)) continue;
        float4 w = S.sa[sq];
        for (int d = 0; d < 4; d++) {
            int nx = qx + NB_X[d], ny = qy + NB_Y[d];
            if (!(S.sd[sidx(nx, ny, x0, y0, 2)] > 0.0f)) continue;
            float2 zn = S.sz[sidx(nx, ny, x0, y0, 1)];
            float wd = f4c(w, d);
            sum.x += wd * zn.x;
            sum.y += wd * zn.y;
        }
    }
    MG<H>::set_R(c, at(C, cx, cy), sum);
}

// Loads the diagonals (ring 2) and couplings (ring 1) of the tile of level l.
template <bool H>
SP_DEV void stage_matrix(const Ctx& c, const Lv& F, int l, int x0, int y0, DownStage& S) {
    const int t = tid2();
    for (int k = t; k < R2 * R2; k += 256) {
        int gx = x0 - 2 + k % R2, gy = y0 - 2 + k / R2;
        S.sd[k] = inside(F, gx, gy) ? MG<H>::D(c, l, at(F, gx, gy)) : 0.0f;
    }
    for (int k = t; k < R1 * R1; k += 256) {
        int gx = x0 - 1 + k % R1, gy = y0 - 1 + k / R1;
        S.sa[k] = inside(F, gx, gy) ? MG<H>::A(c, l, at(F, gx, gy)) : make_float4(0, 0, 0, 0);
    }
}

}  // namespace

// mg_down.comp:36-69 on level l: pre-smoothing from zero and restriction of the residual.
// r_in: the level's right-hand side (level 0: the CG residual buffer of this iteration).
template <bool H>
__global__ __launch_bounds__(256) void k_mg_down(const TankBatch tb, int l, int gated, int rsel) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const float2* __restrict__ r_in = rsel ? c.r2 : c.mgr;
    __shared__ DownStage S;
    if (gated && solve_done(c)) return;
    const Lv F = level(c, l), C = level(c, l + 1);
    if (c.ptile[F.toff + tile_slot()] == 0u) return;
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    stage_matrix<H>(c, F, l, x0, y0, S);
    for (int k = tid2(); k < R2 * R2; k += 256) {
        int gx = x0 - 2 + k % R2, gy = y0 - 2 + k / R2;
        float2 rv = make_float2(0.0f, 0.0f);
        if (inside(F, gx, gy)) rv = l == 0 ? r_in[at(F, gx, gy)] : MG<H>::R(c, at(F, gx, gy));
        S.sr[k] = rv;
    }
    __syncthreads();
    down_compute<H>(c, F, C, x0, y0, S);
}

// cg_update.comp fused with the way down on level 0: x += alpha p, r_out = r_in - alpha q
// (residual buffers alternate between iterations, so a tile may read its neighbours' old
// residual while they write their new one), then mg_down level 0 on r_out.
template <bool H>
__global__ __launch_bounds__(256) void k_update_down0(const TankBatch tb, int par) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    // residual (mgr / r2) and direction (cgp / cgp2) buffers of iteration parity `par`
    const float2* __restrict__ r_in = par ? c.r2 : c.mgr;
    float2* __restrict__ r_out = par ? c.mgr : c.r2;
    const float2* __restrict__ p_new = par ? c.cgp : c.cgp2;
    __shared__ DownStage S;
    if (solve_done(c)) return;
    const Lv F = level(c, 0), C = level(c, 1);
    if (c.ptile[tile_slot()] == 0u) return;
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    const float ax = (float)c.K->alpha[0], ay = (float)c.K->alpha[1];
    stage_matrix<H>(c, F, 0, x0, y0, S);
    for (int k = tid2(); k < R2 * R2; k += 256) {
        int gx = x0 - 2 + k % R2, gy = y0 - 2 + k / R2;
        float2 rv = make_float2(0.0f, 0.0f);
        if (inside(F, gx, gy)) {
            int i = at(F, gx, gy);
            rv = r_in[i];
            if (c.mgd[i] > 0.0f) {
                float2 q = c.cgq[i];
                rv.x -= ax * q.x;
                rv.y -= ay * q.y;
            }
        }
        S.sr[k] = rv;
    }
    {
        const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
        if (inside(F, x, y)) {
            int i = at(F, x, y);
            if (c.mgd[i] > 0.0f) {
                float2 p = p_new[i];
                float2 x1 = c.pres[i];
                x1.x += ax * p.x;
                x1.y += ay * p.y;
                c.pres[i] = x1;
            }
        }
    }
    __syncthreads();
    {
        const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
        if (inside(F, x, y)) r_out[at(F, x, y)] = S.sr[sidx(x, y, x0, y0, 2)];
    }
    down_compute<H>(c, F, C, x0, y0, S);
}

// mg_up.comp:38-75 on level l: coarse-grid correction and the black-then-red
// post-smoothing; on level 0 (reduce >= 0) also the partial sums r.z, r.r.
template <bool H>
__global__ __launch_bounds__(256) void k_mg_up(const TankBatch tb, int l, int gated, int reduce, int nblocks, int rsel) {
    using M = MG<H>;
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const float2* __restrict__ r_in = rsel ? c.r2 : c.mgr;
    __shared__ double4 lds[4];
    __shared__ float sd[R2 * R2];
    __shared__ float2 sx[R2 * R2];
    __shared__ float4 sa[R1 * R1];
    __shared__ float2 sr[R1 * R1];
    __shared__ float2 szc[10 * 10];
    __shared__ float2 sb[R1 * R1];
    if (gated && solve_done(c)) return;
    const Lv F = level(c, l), C = level(c, l + 1);
    const int slot = tile_slot();
    if (c.ptile[F.toff + slot] == 0u) {
        if (reduce >= 0 && tid2() == 0) store_part(&c.part_sums[slot], make_double4(0, 0, 0, 0));
        return;
    }
    const int t = tid2();
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    for (int k = t; k < R2 * R2; k += 256) {
        int gx = x0 - 2 + k % R2, gy = y0 - 2 + k / R2;
        bool in = inside(F, gx, gy);
        sd[k] = in ? M::D(c, l, at(F, gx, gy)) : 0.0f;
        sx[k] = in ? M::X(c, at(F, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    for (int k = t; k < R1 * R1; k += 256) {
        int gx = x0 - 1 + k % R1, gy = y0 - 1 + k / R1;
        bool in = inside(F, gx, gy);
        sa[k] = in ? M::A(c, l, at(F, gx, gy)) : make_float4(0, 0, 0, 0);
        sr[k] = in ? (l == 0 ? r_in[at(F, gx, gy)] : M::R(c, at(F, gx, gy))) : make_float2(0.0f, 0.0f);
    }
    // parents of the ring-2 region: coarse cells x0/2 - 1 .. x0/2 + 8
    const int cx0 = (x0 >> 1) - 1, cy0 = (y0 >> 1) - 1;
    for (int k = t; k < 100; k += 256) {
        int gx = cx0 + k % 10, gy = cy0 + k / 10;
        szc[k] = inside(C, gx, gy) ? M::Z(c, at(C, gx, gy)) : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    // new values of the black cells of the tile and its first ring (mg_up.comp:26-36)
    for (int k = t; k < R1 * R1; k += 256) {
        int gx = x0 - 1 + k % R1, gy = y0 - 1 + k / R1;
        float2 z = make_float2(0.0f, 0.0f);
        float D = sd[sidx(gx, gy, x0, y0, 2)];
        if (D > 0.0f && ((gx + gy) & 1) != 0) {
            float4 w = sa[k];
            float2 sum = sr[k];
            for (int d = 0; d < 4; d++) {
                int mx = gx + NB_X[d], my = gy + NB_Y[d];
                int sm = sidx(mx, my, x0, y0, 2);
                if (!(sd[sm] > 0.0f)) continue;
                float wd = f4c(w, d);
                float2 xm = sx[sm];
                float2 zc = szc[((my >> 1) - cy0) * 10 + ((mx >> 1) - cx0)];
                sum.x += wd * (xm.x + SP_MG_OMEGA * zc.x);
                sum.y += wd * (xm.y + SP_MG_OMEGA * zc.y);
            }
            z = make_float2(sum.x / D, sum.y / D);
        }
        sb[k] = z;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    double4 acc = make_double4(0, 0, 0, 0);
    if (inside(F, x, y)) {
        int i = at(F, x, y);
        int s1 = sidx(x, y, x0, y0, 1);
        float D = sd[sidx(x, y, x0, y0, 2)];
        float2 z = make_float2(0.0f, 0.0f);
        if (D > 0.0f) {
            float2 rv = sr[s1];
            if (((x + y) & 1) == 0) {
                float4 w = sa[s1];
                float2 sum = rv;
                for (int d = 0; d < 4; d++) {
                    int nx = x + NB_X[d], ny = y + NB_Y[d];
                    if (!(sd[sidx(nx, ny, x0, y0, 2)] > 0.0f)) continue;
                    float wd = f4c(w, d);
                    float2 zn = sb[sidx(nx, ny, x0, y0, 1)];
                    sum.x += wd * zn.x;
                    sum.y += wd * zn.y;
                }
                z = make_float2(sum.x / D, sum.y / D);
            } else {
                z = sb[s1];
            }
            acc = make_double4((double)rv.x * (double)z.x, (double)rv.y * (double)z.y, (double)rv.x * (double)rv.x,
                               (double)rv.y * (double)rv.y);
        }
        M::set_Z(c, i, z);
    }
    if (reduce >= 0) {
        double4 s2 = group_sum4(acc, lds);
        if (t == 0) store_part(&c.part_sums[slot], s2);
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

// LDS (dynamic, sized at launch to the hierarchy): per tail cell its couplings (float4),
// iterate and right-hand side (float2) and diagonal; the first level's right-hand side is
// loaded once with the matrix, so that the sweeps never wait on global memory.
template <bool H>
__global__ __launch_bounds__(1024) void k_mg_tail(const TankBatch tb, int gated) {
    using M = MG<H>;
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    extern __shared__ float4 tail_lds[];
    __shared__ int4 sl[SP_MAX_LEVELS];
    if (gated && solve_done(c)) return;
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
    for (int t = bt; t < total; t += 1024) {
        int l = first;  // the level of tail cell t (the storage of the coarse matrices is per level)
        if (H)
            while (l < last && BASE + t >= sl[l + 1].x) l++;
        sd[t] = M::D(c, l, BASE + t);
        sa[t] = M::A(c, l, BASE + t);
        if (t < NF) sr[t] = M::R(c, BASE + t);
    }
    __syncthreads();

    auto rhs_at = [&](int s) -> float2 { return sr[s]; };
    auto sat = [&](int4 L, int x, int y) { return L.x - BASE + y * L.y + x; };
    auto sunk = [&](int4 L, int x, int y) { return x >= 0 && y >= 0 && x < L.y && y < L.z && sd[sat(L, x, y)] > 0.0f; };
    auto sweep = [&](int4 L, int colour, bool zero, int t0, int step) {
        for (int t = t0; t < L.y * L.z; t += step) {
            int x = t % L.y, y = t / L.y;
            int s = L.x - BASE + t;
            float D = sd[s];
            if (D <= 0.0f || ((x + y) & 1) != colour) {
                if (zero) sx[s] = make_float2(0.0f, 0.0f);
                continue;
            }
            float2 sum = rhs_at(s);
            if (!zero) {
                float4 w = sa[s];
                for (int d = 0; d < 4; d++) {
                    int nx = x + NB_X[d], ny = y + NB_Y[d];
                    if (sunk(L, nx, ny)) {
                        float wd = f4c(w, d);
                        float2 xn = sx[sat(L, nx, ny)];
                        sum.x += wd * xn.x;
                        sum.y += wd * xn.y;
                    }
                }
            }
            sx[s] = make_float2(sum.x / D, sum.y / D);
        }
    };
    auto restrict_to = [&](int4 F, int4 C, int t0, int step) {
        for (int t = t0; t < C.y * C.z; t += step) {
            int cx = t % C.y, cy = t / C.y;
            float2 sum = make_float2(0.0f, 0.0f);
            for (int k = 0; k < 4; k++) {
                int x = 2 * cx + (k & 1), y = 2 * cy + (k >> 1);
                if (!sunk(F, x, y)) continue;
                int s = sat(F, x, y);
                float4 w = sa[s];
                float2 az = make_float2(sd[s] * sx[s].x, sd[s] * sx[s].y);
                for (int d = 0; d < 4; d++) {
                    int nx = x + NB_X[d], ny = y + NB_Y[d];
                    if (sunk(F, nx, ny)) {
                        float wd = f4c(w, d);
                        float2 xn = sx[sat(F, nx, ny)];
                        az.x -= wd * xn.x;
                        az.y -= wd * xn.y;
                    }
                }
                float2 r = rhs_at(s);
                sum.x += r.x - az.x;
                sum.y += r.y - az.y;
            }
            sr[C.x - BASE + t] = sum;
        }
    };
    auto post_black = [&](int4 F, int4 C, int t0, int step) {
        for (int t = t0; t < F.y * F.z; t += step) {
            int x = t % F.y, y = t / F.y;
            if (((x + y) & 1) == 0) continue;
            int s = F.x - BASE + t;
            float D = sd[s];
            if (D <= 0.0f) continue;
            float4 w = sa[s];
            float2 sum = rhs_at(s);
            for (int d = 0; d < 4; d++) {
                int nx = x + NB_X[d], ny = y + NB_Y[d];
                if (!sunk(F, nx, ny)) continue;
                float wd = f4c(w, d);
                float2 xn = sx[sat(F, nx, ny)];
                float2 zc = sx[sat(C, nx >> 1, ny >> 1)];
                sum.x += wd * (xn.x + SP_MG_OMEGA * zc.x);
                sum.y += wd * (xn.y + SP_MG_OMEGA * zc.y);
            }
            sx[s] = make_float2(sum.x / D, sum.y / D);
        }
    };

    int ls = first;
    while (ls < last && sl[ls].y * sl[ls].z > SP_TAIL_WARP_CELLS) ls++;
    for (int l = first; l < ls; l++) {
        sweep(sl[l], 0, true, bt, 1024);
        __syncthreads();
        sweep(sl[l], 1, false, bt, 1024);
        __syncthreads();
        restrict_to(sl[l], sl[l + 1], bt, 1024);
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
        post_black(sl[l], sl[l + 1], bt, 1024);
        __syncthreads();
        sweep(sl[l], 0, false, bt, 1024);
        __syncthreads();
    }
    for (int t = bt; t < NF; t += 1024) M::set_Z(c, BASE + t, sx[t]);
}

// The scalar step of the CG (cg_reduce.comp / reduce_tail.glsl modes 0, 1, 2) over the
// partial sums the previous kernel wrote: one 256 work-item group, fixed summation order.
// The reference folds this into the last work-group of the producing kernel; here the
// kernel boundary orders the partials instead, which on a multi-XCD part is far cheaper
// than a device-scope release fence in every producing work-group.
__global__ __launch_bounds__(REDUCE_NT) void k_cg_reduce(const TankBatch tb, int mode, int nblocks, int gated) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ double4 lds[REDUCE_NT / 64];
    if (gated && solve_done(c)) return;
    reduce_scalars(c, mode, nblocks, lds);
}

// Solve statistics at the end of a solve (reduce_tail.glsl:105-116, modes 3 and 9: the
// second solve keeps its own iteration count).
__global__ void k_cg_end(const TankBatch tb, int second) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    sp_solve* K = c.K;
    uint32_t it = K->iters[0] > K->iters[1] ? K->iters[0] : K->iters[1];
    if (second) {
        if (K->iters[0] > K->p2_it_max) K->p2_it_max = K->iters[0];
        if (K->done[0] == 0u) K->p2_unconv += 1u;
        K->p2_iters_last = K->iters[0];
    } else {
        if (it > K->it_max) K->it_max = it;
        if (K->done[0] == 0u) K->unconv += 1u;
        K->iters_last = it;
    }
    float rel = K->bb[0] > 0.0 ? (float)sqrt(K->rr[0] / K->bb[0]) : 0.0f;
    if (rel > K->res_max) K->res_max = rel;
}
