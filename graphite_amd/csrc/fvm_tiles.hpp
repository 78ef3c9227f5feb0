// Slot-parallel tile kernels of the solver's iteration.
//
// A tile is a run of consecutive interior cells whose face slots fit one block (at most
// kTileSlots); one thread evaluates one face slot of one cell, the per-slot terms go to LDS,
// and one thread per cell then accumulates its slots in slot order -- the order of the
// reference's per-cell loops -- so the binary64 results equal the one-thread-per-cell
// kernels bit for bit while the launch has one thread per slot instead of one per cell.
//
// Boundary conditions are fused: a ghost's only neighbour is the cell whose face it lies on,
// so its state (euler_slip_bc / euler_inletoutlet_bc, src/models/euler.cu:412-520) is a
// function of that cell's state and the face normal, evaluated where a kernel needs it.
// Ghost states are therefore never stored during the iteration.
//
// Tiles are mapped to blocks so that each of the 8 XCDs of the MI355X walks a contiguous
// range of tiles (workgroups are dispatched round-robin over XCDs): neighbours in the
// reverse Cuthill-McKee band then hit the XCD's own L2.
#pragma once

#include "fvm_kernels.hpp"

namespace gr {

constexpr int kTileSlots = 248;   // 20 binary64 rows x 248 slots = 39.7 KB LDS: 4 blocks per CU
constexpr int kTileCells = 51;    // one thread per (component, cell) in the per-cell phases
static_assert(5 * kTileCells <= kBlock, "per-cell phases need 5 threads per cell");

// A tile as one 64-byte record (16 words): its first cell c0, its first face slot q0, ncell |
// nslot << 16, then the ncell+1 slot offsets of its cells from q0 as bytes (at most 52, each at
// most kTileSlots).  A block reads everything it needs to start from this one record.
constexpr int kTileRec = 16;
static_assert(kTileCells + 1 <= 4 * (kTileRec - 3) && kTileSlots < 256, "tile record layout");


// Ghost state from the state of its cell, outward normal n (src/models/euler.cu:436-441,
// :473-518).  Same arithmetic, in the same order, as k_euler_bc.
template <typename S>
__device__ inline void ghost_state(const BcRegion &R, const S n[3], const S u1[5], S g[5]) {
    const S r1 = u1[0];
    const S ru1[3] = {u1[1], u1[2], u1[3]};
    const S e1 = u1[4];
    if (R.kind == GR_BC_SLIP) {
        const S un = (n[0] * ru1[0] + n[1] * ru1[1]) + n[2] * ru1[2];
        g[0] = r1;
#pragma unroll
        for (int k = 0; k < 3; ++k) g[1 + k] = ru1[k] - (n[k] * S(2)) * un;
        g[4] = e1;
        return;
    }
    const S u[3] = {fdiv(ru1[0], r1), fdiv(ru1[1], r1), fdiv(ru1[2], r1)};
    const S p = S(0.4) * (e1 - fdiv(S(0.5), r1) * ((ru1[0] * ru1[0] + ru1[1] * ru1[1]) + ru1[2] * ru1[2]));
    const S tt = fdiv(p, S(1) * r1);
    const S mach_i = dsqrt(fdiv((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2], fdiv(S(1.4) * p, r1)));
    const S p_in = (S)R.p, T_in = (S)R.T;
    const S mach[3] = {(S)R.mach[0], (S)R.mach[1], (S)R.mach[2]};
    const S rho_in = fdiv(p_in, S(1) * T_in);
    const S cin = dsqrt(fdiv(S(1.4) * p_in, rho_in));
    if ((u[0] * n[0] + u[1] * n[1]) + u[2] * n[2] < S(0)) {
        if (mach_i > S(1)) {
            const S vin[3] = {mach[0] * cin, mach[1] * cin, mach[2] * cin};
            g[0] = rho_in;
#pragma unroll
            for (int k = 0; k < 3; ++k) g[1 + k] = (mach[k] * rho_in) * cin;
            g[4] = fdiv(p_in, S(0.4)) + (S(0.5) * rho_in) * ((vin[0] * vin[0] + vin[1] * vin[1]) + vin[2] * vin[2]);
        } else {
            const S r0 = fdiv(p, S(1) * T_in);
            const S c = dsqrt(fdiv(S(1.4) * p, r0));
            const S v[3] = {mach[0] * c, mach[1] * c, mach[2] * c};
            g[0] = r0;
#pragma unroll
            for (int k = 0; k < 3; ++k) g[1 + k] = v[k] * r0;
            g[4] = fdiv(p, S(0.4)) + (S(0.5) * r0) * ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        }
    } else {
        if (mach_i > S(1)) {
#pragma unroll
            for (int k = 0; k < 5; ++k) g[k] = u1[k];
        } else {
            const S r0 = fdiv(p_in, S(1) * tt);
            g[0] = r0;
#pragma unroll
            for (int k = 0; k < 3; ++k) g[1 + k] = u[k] * r0;
            g[4] = fdiv(p_in, S(0.4)) + (S(0.5) * r0) * ((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
        }
    }
}

// Index of the table region holding ghost g, or -1 (ghost state then read from memory).
__device__ inline int bc_region(const BcTable &t, uint32_t g) {
    int r = -1;
    for (int k = 0; k < t.n; ++k)
        if (g >= t.r[k].begin && g < t.r[k].end) r = k;
    return r;
}

// Outward unit normal of face f seen from the cell centred at (xi,yi,zi): the reference's
// outer_normal(n_f, fc - x_i) (include/datatypes.cuh:213-215).
template <typename S>
__device__ inline void outward_normal(const MeshView &m, uint32_t f, double xi, double yi, double zi, S n[3]) {
    const double dx = m.fcx[f] - xi, dy = m.fcy[f] - yi, dz = m.fcz[f] - zi;
    const double nfx = m.fnx[f], nfy = m.fny[f], nfz = m.fnz[f];
    const double dn = (nfx * dx + nfy * dy) + nfz * dz;
    const double sg = (double)(dn > 0) - (double)(dn < 0);
    n[0] = (S)(nfx * sg);
    n[1] = (S)(nfy * sg);
    n[2] = (S)(nfz * sg);
}

// Locates the local cell of local slot s among the tile's starts (ncell+1 entries).
__device__ inline int tile_cell_of(const uint32_t *starts, int ncell, uint32_t s) {
    int lo = 0, hi = ncell;   // starts[lo] <= s < starts[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (starts[mid] <= s) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Loads W consecutive values with WB-byte loads (p WB-byte aligned, W values a whole number of
// WB-byte words; WB 16 or 8).
template <typename S, int W, int WB = 16>
__device__ inline void load_vec(const S *__restrict__ p, S (&v)[W]) {
    constexpr int L = WB / (int)sizeof(S);
    static_assert(L >= 1 && W % L == 0, "records are whole load words");
    using word = S __attribute__((ext_vector_type(L)));
    const word *q = reinterpret_cast<const word *>(p);
#pragma unroll
    for (int k = 0; k < W / L; ++k) {
        const word t = q[k];
#pragma unroll
        for (int e = 0; e < L; ++e) v[L * k + e] = t[e];
    }
}

// Gradient and limiter of a cell as one record.  binary64: rows 3c+k of the gradient (component
// c, direction k), then the five limiters -- the reference's operands, 160 contiguous bytes,
// fetched with 16-byte loads from two or three cache lines where twenty separate arrays would
// touch twenty lines.  binary32: the 15 rows, each multiplied by the limiter the reconstruction
// applies to it, and one value of padding (64 bytes): (g.d) l becomes (g l).d, the same
// quantity rounded differently, and a neighbour costs 64 bytes instead of 80.
template <typename S>
constexpr int kGL = sizeof(S) == 8 ? 20 : 16;

// The limiter (component) that multiplies gradient row r in the reconstruction: rho and rho E
// their own; momentum direction k of every component the limiter of component 1 + k (the
// reference's tensor * vector), or the row's own component with the corrected form.
__host__ __device__ constexpr int row_limiter(int r, int corr) {
    return r < 3 ? 0 : r >= 12 ? 4 : corr ? r / 3 : 1 + r % 3;
}
// row_limiter(r, corr) of rows 0..14 packed 3 bits per row (bits 3r..3r+2), for a lookup by shift.
__host__ __device__ constexpr uint64_t row_limiters_packed(int corr) {
    uint64_t v = 0;
    for (int r = 0; r < 15; ++r) v |= (uint64_t)row_limiter(r, corr) << (3 * r);
    return v;
}

// Per-slot geometry of the iteration, one record per face slot q of an interior cell i
// (neighbour j = cell_connects[q], face f = cell_faces[q]), computed once from the binary64
// metrics and cast exactly as the kernels cast them, so reading it instead of gathering face
// and cell metrics leaves every bit of the results unchanged.  The records are read in slot
// order, one contiguous run per tile, where the metrics they replace are gathered by face
// and neighbour id.
//   flux record:     n = outer_normal(n_f, fc - x_i) (3), ds = A_f / V_i, d0 = fc - x_i (3),
//                    d1 = fc - x_j (3): 80 bytes in binary64 (16-byte loads), 40 in binary32
//                    (8-byte loads)
//   gradient record: d = x_j - x_i (3), d0 (3): 48 bytes in binary64 (16-byte loads), 24 in
//                    binary32 (8-byte loads); the kernel forms 1/|d|^2 from d with the
//                    expression that set-up would store, and reads the neighbour from cc
//   ghost record:    n of the ghost's face seen from its cell (3), padding (fused BCs)
constexpr int kFluxRec = 10;
constexpr int kGradRec = 6;
constexpr int kGhostRec = 4;

template <typename S>
__global__ __launch_bounds__(kBlock) void k_slot_geometry(MeshView m, S *flux_rec, S *grad_rec, S *ghost_rec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N) return;
    const double xi = m.cx[i], yi = m.cy[i], zi = m.cz[i];
    for (uint32_t q = m.cfs[i]; q < m.cfs[i + 1]; ++q) {
        const uint32_t f = m.cf[q], j = m.cc[q];
        S n[3];
        outward_normal<S>(m, f, xi, yi, zi, n);
        const double fcx = m.fcx[f], fcy = m.fcy[f], fcz = m.fcz[f];
        const S d0[3] = {(S)(fcx - xi), (S)(fcy - yi), (S)(fcz - zi)};
        S *fr = flux_rec + (size_t)q * kFluxRec;
#pragma unroll
        for (int k = 0; k < 3; ++k) fr[k] = n[k];
        fr[3] = geo_ratio<S>(m.fa[f], m.cv[i]);
#pragma unroll
        for (int k = 0; k < 3; ++k) fr[4 + k] = d0[k];
        fr[7] = (S)(fcx - m.cx[j]);
        fr[8] = (S)(fcy - m.cy[j]);
        fr[9] = (S)(fcz - m.cz[j]);
        const S d[3] = {(S)(m.cx[j] - xi), (S)(m.cy[j] - yi), (S)(m.cz[j] - zi)};
        S *gq = grad_rec + (size_t)q * kGradRec;
#pragma unroll
        for (int k = 0; k < 3; ++k) gq[k] = d[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) gq[3 + k] = d0[k];
        if (j >= m.N) {   // a ghost's only face slot belongs to this cell
            S *g = ghost_rec + (size_t)(j - m.N) * kGhostRec;
#pragma unroll
            for (int k = 0; k < 3; ++k) g[k] = n[k];
            g[3] = S(0);
        }
    }
}

// The solver's state as one record per cell and ghost: rho, rho*u (3), rho*E and one value of
// padding (48 bytes in binary64, read with 16-byte loads; 24 in binary32, 8-byte loads), so a
// neighbour's state is one contiguous run where five arrays would touch five cache lines.
constexpr int kUR = 6;

template <typename S>
__device__ inline void load_state(const S *__restrict__ u, uint32_t i, S v[5]) {
    const S *p = u + (size_t)i * kUR;
    if constexpr (sizeof(S) == 8) {
        S w[kUR];
        load_vec<S, kUR, 16>(p, w);
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] = w[c];
    } else {
        // a 24-byte record starts 8-byte aligned: one 16-byte load (global loads of whole dwords
        // need only dword alignment on this target) and one 4-byte load, two instructions per
        // gathered neighbour where three 8-byte loads would take three
        typedef float f4 __attribute__((ext_vector_type(4)));
        typedef f4 f4a8 __attribute__((aligned(8)));
        const f4a8 a = *reinterpret_cast<const f4a8 *>(p);
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = a[c];
        v[4] = p[4];
    }
}

// Ghost states of buffer u from its interior cells, for regions with a boundary condition (the
// iteration evaluates them on the fly and never stores them; this serves state downloads).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_ghost_states(MeshView m, S *u, const S *__restrict__ ghost_rec, BcTable t) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= m.G) return;
    const int r = bc_region(t, g);
    if (r < 0) return;
    const uint32_t i = m.cc[m.cfs[m.N + g]];
    S ui[5], n[4], ug[5];
    load_vec<S, 4>(ghost_rec + (size_t)g * kGhostRec, n);
#pragma unroll
    for (int c = 0; c < 5; ++c) ui[c] = u[(size_t)i * kUR + c];
    ghost_state<S>(t.r[r], n, ui, ug);
#pragma unroll
    for (int c = 0; c < 5; ++c) u[((size_t)m.N + g) * kUR + c] = ug[c];
}

// Fills n state records with the values v[0..4] (padding zero).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_fill_states(S *u, size_t n, S v0, S v1, S v2, S v3, S v4) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n * kUR) return;
    const int r = (int)(k % kUR);
    u[k] = r == 0 ? v0 : r == 1 ? v1 : r == 2 ? v2 : r == 3 ? v3 : r == 4 ? v4 : S(0);
}

// LSQ matrices in S (6 blocks of stride values): the cast the gradient applies, done once.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_cast_lsq(MeshView m, S *out, size_t stride) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N) return;
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k * stride + i] = (S)m.mat[k * m.mat_stride + i];
}

template <typename S>
struct SolverArgs {
    const S *u;       // current state records (cells; ghosts only for regions without a BC)
    S *un;            // next state records
    S *gl;            // gradient + limiter records of the interior cells, kGL<S> values each
    const S *frec;    // flux records, kFluxRec per interior face slot
    const S *grec;    // gradient records, kGradRec per interior face slot
    const S *gnrm;    // ghost records, kGhostRec per ghost
    const S *lsq;     // LSQ matrices of the interior cells in S, 6 blocks: xx,yy,zz,xy,xz,yz
    size_t lsq_stride;
};

// The reconstruction u + (g.d) l with the gradient rows premultiplied by their limiters
// (G = g l, row_limiter): u + G.d.
template <typename S, bool CORR>
__device__ inline void reconstruct_pre(const S u[5], const S G[15], const S d[3], S &r, S ru[3], S &re) {
    r = u[0] + ((G[0] * d[0] + G[1] * d[1]) + G[2] * d[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        S gd;
        if (CORR) gd = (G[3 * (1 + k)] * d[0] + G[3 * (1 + k) + 1] * d[1]) + G[3 * (1 + k) + 2] * d[2];
        else gd = (G[3 + k] * d[0] + G[6 + k] * d[1]) + G[9 + k] * d[2];
        ru[k] = u[1 + k] + gd;
    }
    re = u[4] + ((G[12] * d[0] + G[13] * d[1]) + G[14] * d[2]);
}

// Reconstructed state at offset d of cell c from its state u and its gradient+limiter record.
template <typename S, bool CORR>
__device__ inline void recon_at(const S *__restrict__ gl, uint32_t c, const S u[5], const S d[3], S &r, S ru[3],
                                S &re) {
    S v[kGL<S>];
    load_vec<S, kGL<S>>(gl + (size_t)c * kGL<S>, v);
    if constexpr (sizeof(S) == 8) reconstruct<S, CORR>(u, v, v + 15, d, r, ru, re);
    else reconstruct_pre<S, CORR>(u, v, d, r, ru, re);   // rows premultiplied by their limiters
}

// Gradient (LSQ) + neighbour min/max + limiter of the five conservative components
// (FvmEulerStepPrepare, src/models/euler.cu:597-647).
//
// In the solver loop (ctl set) block 0 first finishes the previous iteration: the fixed-order
// sum of its flux launch's tile partials (FvmFieldReduceSum, src/fvmops.cu:369-383) and the
// stopping rule (src/models/euler.cu:158-169), which k_finalize would otherwise do in a launch
// of its own; the flux launch that follows reads the stop flag.  Gradients computed by the
// other blocks after the solve has stopped are scratch and never used.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_gradlim_tile(MeshView m, SolverArgs<S> a, BcTable bc,
                                                         const uint32_t *__restrict__ tiles, uint32_t ntiles,
                                                         int kind, int corr, SolverCtl *ctl,
                                                         const double *__restrict__ partials, double *hist) {
    // the tile record is read before the stop flag (nothing is written before the test), so
    // the two loads overlap instead of following each other
    const uint32_t *tr = tiles + (size_t)xcd_tile(blockIdx.x, ntiles) * kTileRec;
    const uint32_t c0 = tr[0], q0 = tr[1];
    const int ncell = (int)(tr[2] & 0xFFFFu), nslot = (int)(tr[2] >> 16);
    const uint8_t *starts = reinterpret_cast<const uint8_t *>(tr + 3);
    // the tile's slot offsets into LDS before any other load is issued: the loads after this
    // need the tile record, whose cache line holds the offsets, so the wait costs nothing, and
    // it does not drain them (loads complete in order, so a wait issued after them would)
    __shared__ uint32_t s_start[kTileCells + 1];
    if (threadIdx.x <= (unsigned)ncell) s_start[threadIdx.x] = starts[threadIdx.x];
    if (ctl) {
        __shared__ int s_stop;
        if (blockIdx.x == 0 && ctl->pending) {   // uniform across the block
            const double tot = block_sum_partials(partials, ntiles);
            if (threadIdx.x == 0) {
                record_iteration(ctl, hist, tot);
                s_stop = ctl->stop;
            }
            __syncthreads();
            if (s_stop) return;
        } else if (ctl->stop) {
            return;
        }
    }
    // LDS rows per slot column.  Binary32 (kTerms): 0..14 the slot's gradient terms
    // d*(u_j-u_i)/|d|^2, component c direction k in row 3c+k, and 15..19 u_j.  Binary64: 0..4
    // u_j - u_i, 5..7 d = x_j - x_i, 8 1/|d|^2, 9..13 u_j -- 27.8 KB instead of 39.7 KB, five
    // blocks per CU instead of four; phase B forms the terms itself.  After phase B the cell's
    // first slot columns sb.. hold its gradient (grad_at), minimum and maximum (min_at, max_at);
    // after phase D its limiters sit in row kUj+c of column sb.
    constexpr bool kTerms = sizeof(S) == 4;
    constexpr int kUj = kTerms ? 15 : 9;
    __shared__ S s_buf[kUj + 5][kTileSlots];
    auto grad_at = [&](int c, int k, int sb) -> S & { return kTerms ? s_buf[3 * c + k][sb] : s_buf[c][sb + k]; };
    auto min_at = [&](int c, int sb) -> S & {
        return kTerms ? s_buf[c][sb + 1] : s_buf[5 + (2 * c) % 4][sb + (2 * c) / 4];
    };
    auto max_at = [&](int c, int sb) -> S & {
        return kTerms ? s_buf[5 + c][sb + 1] : s_buf[5 + (2 * c + 1) % 4][sb + (2 * c + 1) / 4];
    };
    const int t = threadIdx.x;
    // loads of the per-(component, cell) phase B (the cell's LSQ matrix and value), issued first
    // so that their latency hides under the slot gathers of phase A
    const bool in_b = t < 5 * ncell;
    const int bc_c = in_b ? t / ncell : 0, bc_l = in_b ? t - bc_c * ncell : 0;
    S bmat[6], bu = S(0);
#pragma unroll
    for (int k = 0; k < 6; ++k) bmat[k] = S(0);
    if (in_b) {
        const uint32_t ic = c0 + bc_l;
#pragma unroll
        for (int k = 0; k < 6; ++k) bmat[k] = a.lsq[k * a.lsq_stride + ic];
        bu = a.u[(size_t)ic * kUR + bc_c];
    }
    // What a slot reads by its own index (its record, the neighbour's state or the ghost's face
    // normal): in binary64 it is loaded before the barrier that the slot -> cell lookup needs,
    // so that its latency runs alongside the lookup's (wing: 164 -> 158 us); in binary32 that
    // was slower (0.1745 -> 0.1769 ms per iteration), and the loads come right after the barrier,
    // before the lookup and the cell's own state (0.1679 -> 0.1660 ms against after them: loads
    // complete in order, so the neighbour's state, which waits for its index, no longer waits
    // for the cell's state as well).
    constexpr bool kEarly = sizeof(S) == 8;
    constexpr int kGradWB = sizeof(S) == 8 ? 16 : 8;
    S dw[kGradRec], uj[5], gn[4];   // dw: d = x_j - x_i, fc - x_i (phase C)
    int r = -1;                     // the ghost neighbour's BC region (-1: state from the array)
    if constexpr (kEarly) {
        if (t < nslot) {
            const uint32_t j = m.cc[q0 + t];
            load_vec<S, kGradRec, kGradWB>(a.grec + (size_t)(q0 + t) * kGradRec, dw);
            if (j >= m.N) r = bc_region(bc, j - m.N);
            if (r < 0) load_state<S>(a.u, j, uj);
            else load_vec<S, 4>(a.gnrm + (size_t)(j - m.N) * kGhostRec, gn);
        }
    }
    __syncthreads();

    // A: per slot u_j and either its gradient terms or u_j - u_i with the slot's offset and weight
    int lc = 0;
    S ui[5], d0[3] = {S(0), S(0), S(0)};
    if (t < nslot) {
        if constexpr (!kEarly) {   // the neighbour's loads first: they need no lookup
            const uint32_t j = m.cc[q0 + t];
            load_vec<S, kGradRec, kGradWB>(a.grec + (size_t)(q0 + t) * kGradRec, dw);
            r = j >= m.N ? bc_region(bc, j - m.N) : -1;
            if (r < 0) load_state<S>(a.u, j, uj);
            else load_vec<S, 4>(a.gnrm + (size_t)(j - m.N) * kGhostRec, gn);
        }
        lc = tile_cell_of(s_start, ncell, (uint32_t)t);
        const uint32_t i = c0 + lc;
        load_state<S>(a.u, i, ui);
        if (r >= 0) ghost_state<S>(bc.r[r], gn, ui, uj);
        d0[0] = dw[3];
        d0[1] = dw[4];
        d0[2] = dw[5];
        // the LSQ weight 1/|d|^2, products summed in x, y, z order
        const S w2 = fdiv(S(1), (dw[0] * dw[0] + dw[1] * dw[1]) + dw[2] * dw[2]);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const S du = uj[c] - ui[c];
            if constexpr (kTerms) {
#pragma unroll
                for (int k = 0; k < 3; ++k) s_buf[3 * c + k][t] = (dw[k] * du) * w2;
            } else {
                s_buf[c][t] = du;
            }
            s_buf[kUj + c][t] = uj[c];
        }
        if constexpr (!kTerms) {
#pragma unroll
            for (int k = 0; k < 3; ++k) s_buf[5 + k][t] = dw[k];
            s_buf[8][t] = w2;
        }
    }
    __syncthreads();

    // B: one thread per (component, cell): the terms d*(u_j-u_i)/|d|^2 summed in slot order, LSQ
    // apply, min/max.  The results go to rows other threads of the cell still read, hence the
    // barrier between the reads and the writes.
    S bg[3] = {S(0), S(0), S(0)}, bmin = S(0), bmax = S(0);
    int bsb = 0;
    if (in_b) {
        bsb = (int)s_start[bc_l];
        const int se = (int)s_start[bc_l + 1];
        S g3[3] = {S(0), S(0), S(0)};
        bmin = bmax = bu;
        for (int s = bsb; s < se; ++s) {
            if constexpr (kTerms) {
#pragma unroll
                for (int k = 0; k < 3; ++k) g3[k] += s_buf[3 * bc_c + k][s];
            } else {
                const S du = s_buf[bc_c][s], w2 = s_buf[8][s];
#pragma unroll
                for (int k = 0; k < 3; ++k) g3[k] += (s_buf[5 + k][s] * du) * w2;
            }
            bmin = fmin(bmin, s_buf[kUj + bc_c][s]);
            bmax = fmax(bmax, s_buf[kUj + bc_c][s]);
        }
        // M * g with M symmetric (xx,yy,zz,xy,xz,yz), as lsq_apply
        bg[0] = (bmat[0] * g3[0] + bmat[3] * g3[1]) + bmat[4] * g3[2];
        bg[1] = (bmat[3] * g3[0] + bmat[1] * g3[1]) + bmat[5] * g3[2];
        bg[2] = (bmat[4] * g3[0] + bmat[5] * g3[1]) + bmat[2] * g3[2];
    }
    __syncthreads();
    if (in_b) {
#pragma unroll
        for (int k = 0; k < 3; ++k) grad_at(bc_c, k, bsb) = bg[k];
        min_at(bc_c, bsb) = bmin;
        max_at(bc_c, bsb) = bmax;
    }
    __syncthreads();

    // C: per-slot limiter of each component at this slot's face centre.  Reads rows below kUj
    // of the cell's first slot columns, writes rows kUj.. of its own column (u_j, consumed in
    // B): disjoint, so no barrier is needed between the reads and the writes.
    if (t < nslot) {
        const int sb = (int)s_start[lc];
#pragma unroll
        for (int c = 0; c < 5; ++c)
            s_buf[kUj + c][t] = limiter_face<S>(ui[c], grad_at(c, 0, sb), grad_at(c, 1, sb), grad_at(c, 2, sb),
                                                min_at(c, sb), max_at(c, sb), d0[0], d0[1], d0[2], kind);
    }
    __syncthreads();

    // D: one thread per (component, cell): minimum over the cell's faces, starting from 2.  Row
    // kUj+c of the cell's slots is read by this thread alone, so the limiter can go straight to
    // the cell's first slot column.
    if (in_b) {
        const int se = (int)s_start[bc_l + 1];
        S l = S(2);
        for (int s = bsb; s < se; ++s) l = fmin(l, s_buf[kUj + bc_c][s]);
        s_buf[kUj + bc_c][bsb] = l;
    }
    __syncthreads();
    // E: the tile's gradient + limiter records, contiguous in memory, written coalesced: in
    // binary32 the limited rows as 16-byte words, in binary64 value by value (16-byte words
    // measured 1% slower there)
    if constexpr (sizeof(S) == 4) {
        constexpr int WG = kGL<S> / 4;
        using word = float __attribute__((ext_vector_type(4)));
        word *dst = reinterpret_cast<word *>(a.gl + (size_t)c0 * kGL<S>);
        const uint64_t lrows = corr ? row_limiters_packed(1) : row_limiters_packed(0);
        for (int k = t; k < WG * ncell; k += blockDim.x) {
            const int cl = k / WG, r0 = (k - cl * WG) * 4, col = (int)s_start[cl];
            word v;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = r0 + e;   // gradient row r (binary32 layout: LDS row r), padding at 15
                const int lr = (int)((lrows >> (3 * r)) & 7u);
                v[e] = r < 15 ? s_buf[r][col] * s_buf[kUj + lr][col] : S(0);
            }
            dst[k] = v;
        }
    } else {
        for (int k = t; k < kGL<S> * ncell; k += blockDim.x) {
            const int cl = k / kGL<S>, r = k - cl * kGL<S>, col = (int)s_start[cl];
            a.gl[(size_t)c0 * kGL<S> + k] = r < 15 ? grad_at(r / 3, r % 3, col) : s_buf[kUj + (r - 15)][col];
        }
    }
}

// Flux + local-time-step update (fvmEulerSteadyOp + euler_steady_step,
// src/models/euler.cu:236-381), Jacobi, with V*|R|^2 block partials.
template <typename S, bool CORR>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(S) == 8 ? 4 : 8, 8))) void k_flux_tile(MeshView m, SolverArgs<S> a, BcTable bc,
                                                      const uint32_t *__restrict__ tiles, uint32_t ntiles, double cfl,
                                                      double *partials, SolverCtl *ctl) {
    // the tile record is read before the stop flag (nothing is written before the test), so
    // the two loads overlap instead of following each other
    const uint32_t *tr = tiles + (size_t)xcd_tile(blockIdx.x, ntiles) * kTileRec;
    const uint32_t c0 = tr[0], q0 = tr[1];
    const int ncell = (int)(tr[2] & 0xFFFFu), nslot = (int)(tr[2] >> 16);
    const uint8_t *starts = reinterpret_cast<const uint8_t *>(tr + 3);
    // the tile's slot offsets into LDS before any other load is issued: the loads after this
    // need the tile record, whose cache line holds the offsets, so the wait costs nothing, and
    // it does not drain them (loads complete in order, so a wait issued after them would)
    __shared__ uint32_t s_start[kTileCells + 1];
    if (threadIdx.x <= (unsigned)ncell) s_start[threadIdx.x] = starts[threadIdx.x];
    if (ctl) {
        if (ctl->stop) return;
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl->pending = 1;   // partials for the next launch to sum
    }
    __shared__ S s_c[5][kTileSlots];   // central flux term ((F0+F1)*ds)*0.5
    __shared__ S s_d[5][kTileSlots];   // dissipation term ((eig*(U1-U0))*0.5)*ds
    __shared__ S s_e[kTileSlots];      // eig*ds
    const int t = threadIdx.x;
    const bool in_b = t < 5 * ncell;   // one thread per (component, cell) in the update
    const int cc_ = in_b ? t / ncell : 0, cl = in_b ? t - cc_ * ncell : 0;
    // What a slot reads by its own index (neighbour, record, the neighbour's state) is loaded
    // before the barrier that the slot -> cell lookup needs, so that its latency runs alongside
    // the lookup's (wing: binary32 95.3 -> 90.3 us, binary64 221.6 -> 219.0 us; the neighbour's
    // gradient record as well left binary32 no faster, 0.1675 -> 0.1680 ms per iteration)
    uint32_t j = 0;
    int jr = -1;   // the ghost neighbour's BC region (-1: state read from the array)
    S geo[kFluxRec], uj[5];
    if (t < nslot) {
        j = m.cc[q0 + t];
        load_vec<S, kFluxRec, 2 * sizeof(S) == 16 ? 16 : 8>(a.frec + (size_t)(q0 + t) * kFluxRec, geo);
        if (j >= m.N) jr = bc_region(bc, j - m.N);
        if (jr < 0) load_state<S>(a.u, j, uj);
    }
    __syncthreads();

    if (t < nslot) {
        const int lc = tile_cell_of(s_start, ncell, (uint32_t)t);
        const uint32_t i = c0 + lc;
        const S n[3] = {geo[0], geo[1], geo[2]};
        const S ds = geo[3];
        const S d0[3] = {geo[4], geo[5], geo[6]};
        const S d1[3] = {geo[7], geo[8], geo[9]};
        S ui[5];
        load_state<S>(a.u, i, ui);
        S r0, ru0[3], e0, r1, ru1[3], e1;
        recon_at<S, CORR>(a.gl, i, ui, d0, r0, ru0, e0);
        if (j < m.N) {
            recon_at<S, CORR>(a.gl, j, uj, d1, r1, ru1, e1);
        } else {   // ghost: its state from the BC, gradient and limiter of cell i
            S ug[5];
            if (jr < 0) {
#pragma unroll
                for (int c = 0; c < 5; ++c) ug[c] = uj[c];
            } else {
                ghost_state<S>(bc.r[jr], n, ui, ug);
            }
            recon_at<S, CORR>(a.gl, i, ug, d1, r1, ru1, e1);   // cell i's gradient and limiter
        }
        const S p0 = S(0.4) * (e0 - fdiv(S(0.5), r0) * ((ru0[0] * ru0[0] + ru0[1] * ru0[1]) + ru0[2] * ru0[2]));
        const S p1 = S(0.4) * (e1 - fdiv(S(0.5), r1) * ((ru1[0] * ru1[0] + ru1[1] * ru1[1]) + ru1[2] * ru1[2]));
        const S V0 = fdiv((ru0[0] * n[0] + ru0[1] * n[1]) + ru0[2] * n[2], r0);
        const S V1 = fdiv((ru1[0] * n[0] + ru1[1] * n[1]) + ru1[2] * n[2], r1);
        s_c[0][t] = (((V0 * r0) + (V1 * r1)) * ds) * S(0.5);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const S fu0 = ru0[k] * V0 + n[k] * p0;
            const S fu1 = ru1[k] * V1 + n[k] * p1;
            s_c[1 + k][t] = ((fu0 + fu1) * ds) * S(0.5);
        }
        s_c[4][t] = (((V0 * (e0 + p0)) + (V1 * (e1 + p1))) * ds) * S(0.5);
        const S v0[3] = {fdiv(ru0[0], r0), fdiv(ru0[1], r0), fdiv(ru0[2], r0)};
        const S v1[3] = {fdiv(ru1[0], r1), fdiv(ru1[1], r1), fdiv(ru1[2], r1)};
        const S cs0 = dsqrt(fmax(S(1e-12), fmin(S(1e30), fdiv(S(1.4) * p0, r0))));
        const S cs1 = dsqrt(fmax(S(1e-12), fmin(S(1e30), fdiv(S(1.4) * p1, r1))));
        const S eig = fmax(dsqrt((v0[0] * v0[0] + v0[1] * v0[1]) + v0[2] * v0[2]) + cs0,
                           dsqrt((v1[0] * v1[0] + v1[1] * v1[1]) + v1[2] * v1[2]) + cs1);
        s_d[0][t] = ((eig * (r1 - r0)) * S(0.5)) * ds;
#pragma unroll
        for (int k = 0; k < 3; ++k) s_d[1 + k][t] = (((ru1[k] - ru0[k]) * eig) * S(0.5)) * ds;
        s_d[4][t] = ((eig * (e1 - e0)) * S(0.5)) * ds;
        s_e[t] = eig * ds;
    }
    __syncthreads();

    // one thread per (component, cell): accumulate the slots in order, update; the cell's
    // residual components (rows of s_d) and new state (rows of s_c) meet in LDS, free after
    // the reads, for V*|R|^2 and for whole-record state writes
    S acc = 0, unew = 0;
    if (in_b) {
        const int sb = (int)s_start[cl], se = (int)s_start[cl + 1];
        S dts = 0;
        for (int s = sb; s < se; ++s) {
            acc = (acc + s_c[cc_][s]) - s_d[cc_][s];
            dts += s_e[s];
        }
        const S dt = fdiv((S)cfl, dts);
        unew = a.u[(size_t)(c0 + cl) * kUR + cc_] + (-acc) * dt;
    }
    __syncthreads();
    if (in_b) {
        s_d[cc_][cl] = acc;
        s_c[cc_][cl] = unew;
    }
    __syncthreads();
    // the tile's new state records, contiguous in memory, written coalesced
    for (int k = t; k < kUR * ncell; k += blockDim.x) {
        const int r = k % kUR;
        a.un[(size_t)c0 * kUR + k] = r < 5 ? s_c[r][k / kUR] : S(0);
    }
    double e_loc = 0;
    if (t < ncell) {
        const double r = s_d[0][t], u0 = s_d[1][t], u1 = s_d[2][t], u2 = s_d[3][t], e = s_d[4][t];
        e_loc = ((r * r + ((u0 * u0 + u1 * u1) + u2 * u2)) + e * e) * m.cv[c0 + t];
    }
    const double bs = block_sum(e_loc);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

}  // namespace gr
