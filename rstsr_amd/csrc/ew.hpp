// Shared strided elementwise machinery: the launch plan over up to three
// operands (output c, inputs a and b), the generic / row / LDS-tiled kernels and
// the host-side planner. A body type supplies the element loads (la, lb) and the
// store (st); the kernels only do index arithmetic.
//
// Replaces the layout-iterator loops of rstsr-native-impl/src/cpu_rayon/
// op_with_func.rs:13-390. Results are independent of iteration order, so the
// planner is free to pick the order that coalesces HBM traffic:
//   1. drop size-1 axes, sort axes by the output's |stride| (innermost first)
//      and merge axes that are contiguous for every operand;
//   2. if an input's unit-stride axis differs from the output's, run the
//      LDS-tiled kernel: tiles read along the input's unit axis and written
//      along the output's unit axis (both sides coalesced);
//   3. else if the innermost extent is long, run the row kernel (one block per
//      row chunk, index math once per block);
//   4. else the generic kernel (per-element fast-division unravel).
#pragma once
#include "common.hpp"
#include "registry.hpp"

#include <algorithm>
#include <array>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

namespace rstsr {
namespace hip {

// absent operand
struct None {};
/* =============================== plan =============================== */

struct EwPlan {
    int nd = 0;
    uint32_t shape[MAXD];
    FastDiv fd[MAXD];
    int64_t st[3][MAXD];  // element strides of c, a, b
    int64_t off[3];       // element offsets of c, a, b
};

// The index math of one row/tile/element: offsets of (c, a, b) for a linear
// index over dims [d0, nd).
__device__ __forceinline__ void unravel3(const EwPlan& p, int d0, uint32_t lin, int64_t& oc, int64_t& oa, int64_t& ob) {
    oc = p.off[0];
    oa = p.off[1];
    ob = p.off[2];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
        if (d < d0 || d >= p.nd) continue;
        uint32_t q = fdiv(lin, p.fd[d]);
        uint32_t r = lin - q * p.shape[d];
        oc += (int64_t)r * p.st[0][d];
        oa += (int64_t)r * p.st[1][d];
        ob += (int64_t)r * p.st[2][d];
        lin = q;
    }
}

constexpr int EW_BLOCK = 256;
constexpr int EW_VEC = 4;

// generic: one linear element index per (thread, v), unravel per element
template <class Body>
__global__ void __launch_bounds__(EW_BLOCK) k_ew_generic(Body body, EwPlan p, uint32_t total) {
    using TA = typename Body::TA;
    using TB = typename Body::TB;
    uint32_t base = blockIdx.x * (EW_BLOCK * EW_VEC) + threadIdx.x;
    TA va[EW_VEC];
    TB vb[EW_VEC];
    int64_t oc[EW_VEC];
#pragma unroll
    for (int v = 0; v < EW_VEC; ++v) {
        uint32_t i = base + v * EW_BLOCK;
        if (i < total) {
            int64_t oa, ob;
            unravel3(p, 0, i, oc[v], oa, ob);
            va[v] = body.la(oa);
            vb[v] = body.lb(ob);
        }
    }
#pragma unroll
    for (int v = 0; v < EW_VEC; ++v) {
        uint32_t i = base + v * EW_BLOCK;
        if (i < total) body.st(oc[v], va[v], vb[v]);
    }
}

// rows: dim 0 is a long row; blockIdx -> (row, chunk); unravel once per block.
// NTM: bit0 / bit1 streaming loads of A / B (operands read once), bit2 streaming
// stores (large outputs).
template <class Body, int NTM>
__global__ void __launch_bounds__(EW_BLOCK) k_ew_rows(Body body, EwPlan p, uint32_t nchunk, FastDiv fchunk) {
    using TA = typename Body::TA;
    using TB = typename Body::TB;
    uint32_t row = fdiv(blockIdx.x, fchunk);
    uint32_t chunk = blockIdx.x - row * nchunk;
    int64_t rc, ra, rb;
    unravel3(p, 1, row, rc, ra, rb);
    const uint32_t L = p.shape[0];
    const int64_t sc = p.st[0][0], sa = p.st[1][0], sb = p.st[2][0];
    uint32_t base = chunk * (EW_BLOCK * EW_VEC) + threadIdx.x;
    TA va[EW_VEC];
    TB vb[EW_VEC];
#pragma unroll
    for (int v = 0; v < EW_VEC; ++v) {
        uint32_t i = base + v * EW_BLOCK;
        if (i < L) {
            va[v] = body.template la<(NTM & 1) != 0>(ra + (int64_t)i * sa);
            vb[v] = body.template lb<(NTM & 2) != 0>(rb + (int64_t)i * sb);
        }
    }
#pragma unroll
    for (int v = 0; v < EW_VEC; ++v) {
        uint32_t i = base + v * EW_BLOCK;
        if (i < L) body.template st<(NTM & 4) != 0>(rc + (int64_t)i * sc, va[v], vb[v]);
    }
}

// tiled transpose: dims 0 (= dc, output unit axis) and DI (input unit axis)
// form a TI (along DI) x TC (along dc) tile; remaining dims are unravelled from
// the block index. Operands flagged in STAGE (bit0 = A, bit1 = B) are read
// along DI into LDS. Each thread keeps one tile column (phase 1) / one tile
// row position (phase 2) fixed, so per-element addressing is a constant stride
// added to a per-thread base; interior tiles skip the bounds tests; an operand
// that is broadcast along DI is loaded once per thread in phase 2.
// ONE: at most one axis besides 0 and di (od, or -1 for none), so the outer
// offset is one product instead of an unravel over every axis.
// NTM: bit0 / bit1 streaming loads of the staged A / B in full tiles, bit2
// streaming stores.
template <class Body, int TI, int TC, int STAGE, bool ONE, int NTM>
__global__ void __launch_bounds__(EW_BLOCK) k_ew_tiled(Body body, EwPlan p, int di, int od, uint32_t nti,
                                                       uint32_t ntc, FastDiv fti, FastDiv ftc, int xmap) {
    using TA = typename Body::TA;
    using TB = typename Body::TB;
    constexpr bool SA = (STAGE & 1) != 0;
    constexpr bool SB = (STAGE & 2) != 0;
    constexpr int ROWS1 = EW_BLOCK / TI, PER1 = TC / ROWS1;  // phase 1: TI lanes along DI
    constexpr int ROWS2 = EW_BLOCK / TC, PER2 = TI / ROWS2;  // phase 2: TC lanes along dc
    __shared__ TA lds_a[SA ? TC : 1][SA ? TI + 1 : 1];
    __shared__ TB lds_b[SB ? TC : 1][SB ? TI + 1 : 1];

    // tile order (autotuned per problem class): 0 output-axis tiles fastest,
    // 1 the same in one contiguous range per XCD, 2 input-axis tiles fastest
    uint32_t b = xmap == 1 ? xcd_chunk(blockIdx.x, gridDim.x) : blockIdx.x;
    uint32_t tc, ti, outer;
    if (xmap == 2) {
        uint32_t q1 = fdiv(b, fti);
        ti = b - q1 * nti;
        outer = fdiv(q1, ftc);
        tc = q1 - outer * ntc;
    } else {
        uint32_t q1 = fdiv(b, ftc);
        tc = b - q1 * ntc;
        outer = fdiv(q1, fti);
        ti = q1 - outer * nti;
    }

    // outer offsets: dims other than 0 and di, in order
    int64_t oc = p.off[0], oa = p.off[1], ob = p.off[2];
    if constexpr (ONE) {
        if (od >= 0) {
            oc += (int64_t)outer * p.st[0][od];
            oa += (int64_t)outer * p.st[1][od];
            ob += (int64_t)outer * p.st[2][od];
        }
    } else {
        uint32_t lin = outer;
#pragma unroll
        for (int d = 1; d < MAXD; ++d) {
            if (d >= p.nd || d == di) continue;
            uint32_t q = fdiv(lin, p.fd[d]);
            uint32_t r = lin - q * p.shape[d];
            oc += (int64_t)r * p.st[0][d];
            oa += (int64_t)r * p.st[1][d];
            ob += (int64_t)r * p.st[2][d];
            lin = q;
        }
    }
    const uint32_t Nc = p.shape[0], Ni = p.shape[di];
    const uint32_t c0 = tc * TC, i0 = ti * TI;
    const int64_t sc_c = p.st[0][0], sc_i = p.st[0][di];
    const int64_t sa_c = p.st[1][0], sa_i = p.st[1][di];
    const int64_t sb_c = p.st[2][0], sb_i = p.st[2][di];
    const bool full = (i0 + TI <= Ni) && (c0 + TC <= Nc);
    const int tid = threadIdx.x;

    if constexpr (SA || SB) {
        // phase 1: ii = tid % TI fixed, cc = k*ROWS1 + tid/TI -> coalesced along the input axis
        const int ii = tid % TI, cb = tid / TI;
        const int64_t pa = oa + (int64_t)(i0 + ii) * sa_i + (int64_t)(c0 + cb) * sa_c, da = (int64_t)ROWS1 * sa_c;
        const int64_t pb = ob + (int64_t)(i0 + ii) * sb_i + (int64_t)(c0 + cb) * sb_c, db = (int64_t)ROWS1 * sb_c;
        if (full) {
            // offsets advance by a uniform stride: one 64-bit add per load
            int64_t ua = pa, ub = pb;
#pragma unroll
            for (int k = 0; k < PER1; ++k) {
                if constexpr (SA) lds_a[k * ROWS1 + cb][ii] = body.template la<(NTM & 1) != 0>(ua);
                if constexpr (SB) lds_b[k * ROWS1 + cb][ii] = body.template lb<(NTM & 2) != 0>(ub);
                ua += da;
                ub += db;
            }
        } else {
#pragma unroll
            for (int k = 0; k < PER1; ++k) {
                if (i0 + ii < Ni && c0 + k * ROWS1 + cb < Nc) {
                    if constexpr (SA) lds_a[k * ROWS1 + cb][ii] = body.la(pa + k * da);
                    if constexpr (SB) lds_b[k * ROWS1 + cb][ii] = body.lb(pb + k * db);
                }
            }
        }
        __syncthreads();
    }
    // phase 2: cc = tid % TC fixed, ii = k*ROWS2 + tid/TC -> coalesced along the output axis
    const int cc = tid % TC, ib = tid / TC;
    const int64_t pc = oc + (int64_t)(c0 + cc) * sc_c + (int64_t)(i0 + ib) * sc_i, dc2 = (int64_t)ROWS2 * sc_i;
    const int64_t qa = oa + (int64_t)(c0 + cc) * sa_c + (int64_t)(i0 + ib) * sa_i, da2 = (int64_t)ROWS2 * sa_i;
    const int64_t qb = ob + (int64_t)(c0 + cc) * sb_c + (int64_t)(i0 + ib) * sb_i, db2 = (int64_t)ROWS2 * sb_i;
    const bool col_ok = c0 + cc < Nc;
    auto run = [&](auto hoist_b) {
        constexpr bool HB = decltype(hoist_b)::value;
        TB yb{};
        if constexpr (HB) {
            if (col_ok) yb = body.lb(qb);
        }
        auto one = [&](int k) {
            TA x;
            TB y;
            if constexpr (SA) x = lds_a[cc][k * ROWS2 + ib];
            else x = body.la(qa + k * da2);
            if constexpr (SB) y = lds_b[cc][k * ROWS2 + ib];
            else if constexpr (HB) y = yb;
            else y = body.lb(qb + k * db2);
            body.template st<(NTM & 4) != 0>(pc + k * dc2, x, y);
        };
        if (full) {
            int64_t uc = pc, ua = qa, ub = qb;
#pragma unroll
            for (int k = 0; k < PER2; ++k) {
                TA x;
                TB y;
                if constexpr (SA) x = lds_a[cc][k * ROWS2 + ib];
                else x = body.la(ua);
                if constexpr (SB) y = lds_b[cc][k * ROWS2 + ib];
                else if constexpr (HB) y = yb;
                else y = body.lb(ub);
                body.template st<(NTM & 4) != 0>(uc, x, y);
                uc += dc2;
                ua += da2;
                ub += db2;
            }
        } else {
#pragma unroll
            for (int k = 0; k < PER2; ++k)
                if (col_ok && i0 + k * ROWS2 + ib < Ni) one(k);
        }
    };
    if (!SB && sb_i == 0) run(std::true_type{});
    else run(std::false_type{});
}

/* =============================== host planning =============================== */

struct Dims {
    int nd = 0;  // -1: empty (no elements)
    std::vector<uint64_t> shape;
    std::vector<std::array<int64_t, 3>> st;  // per dim, strides of c, a, b
    int64_t off[3] = {0, 0, 0};
};

// Collapse the common shape of up to three operands (missing operands carry
// stride 0). Order: output |stride| ascending; merge axes contiguous in all.
static Dims collapse3(const Shape& shape, const Stride* s[3], const size_t off[3]) {
    Dims d;
    for (int k = 0; k < 3; ++k) d.off[k] = (int64_t)off[k];
    std::vector<size_t> ax;
    for (size_t i = 0; i < shape.size(); ++i) {
        if (shape[i] == 0) {
            d.nd = -1;
            return d;
        }
        if (shape[i] != 1) ax.push_back(i);
    }
    auto S = [&](int k, size_t i) -> int64_t { return s[k] ? (int64_t)(*s[k])[i] : 0; };
    std::stable_sort(ax.begin(), ax.end(), [&](size_t x, size_t y) {
        for (int k = 0; k < 3; ++k) {
            int64_t ax_ = std::llabs(S(k, x)), ay = std::llabs(S(k, y));
            if (ax_ != ay) return ax_ < ay;
        }
        return false;
    });
    for (size_t i : ax) {
        if (!d.shape.empty()) {
            auto& ps = d.st.back();
            uint64_t pn = d.shape.back();
            bool ok = true;
            for (int k = 0; k < 3; ++k)
                if (S(k, i) != ps[k] * (int64_t)pn) ok = false;
            if (ok) {
                d.shape.back() *= shape[i];
                continue;
            }
        }
        d.shape.push_back(shape[i]);
        d.st.push_back({S(0, i), S(1, i), S(2, i)});
    }
    d.nd = (int)d.shape.size();
    return d;
}

static EwPlan to_plan(const Dims& d) {
    EwPlan p;
    p.nd = d.nd;
    if (d.nd > MAXD)
        RSTSR_RAISE(UnImplemented, "layouts with more than 8 non-mergeable axes are not supported by the HIP device");
    for (int i = 0; i < MAXD; ++i) {
        p.shape[i] = 1;
        p.fd[i] = FastDiv(1);
        for (int k = 0; k < 3; ++k) p.st[k][i] = 0;
    }
    for (int i = 0; i < d.nd; ++i) {
        p.shape[i] = (uint32_t)d.shape[i];
        p.fd[i] = FastDiv((uint32_t)d.shape[i]);
        for (int k = 0; k < 3; ++k) p.st[k][i] = d.st[i][k];
    }
    for (int k = 0; k < 3; ++k) p.off[k] = d.off[k];
    return p;
}

enum class EwKind { Generic, Rows, Tiled };

struct EwChoice {
    EwKind kind = EwKind::Generic;
    int di = -1;
    int stage = 0;  // bit0 A, bit1 B
};

// has_a / has_b: operand present as an array (not scalar / absent)
static EwChoice choose(const Dims& d, bool has_a, bool has_b) {
    EwChoice ch;
    uint64_t total = 1;
    for (auto n : d.shape) total *= n;
    if (d.nd <= 0 || total < 4096) return ch;
    // input-preferred axis: a unit-stride axis of some input that is not the
    // output's innermost axis
    int best_di = -1;
    for (int k = 1; k <= 2; ++k) {
        if ((k == 1 && !has_a) || (k == 2 && !has_b)) continue;
        if (std::llabs(d.st[0][k]) <= 1) continue;  // already contiguous along dim 0
        for (int i = 1; i < d.nd; ++i) {
            if (std::llabs(d.st[i][k]) == 1) {
                if (best_di < 0) best_di = i;
                break;
            }
        }
    }
    if (best_di > 0 && d.shape[0] >= 16 && d.shape[best_di] >= 16) {
        ch.kind = EwKind::Tiled;
        ch.di = best_di;
        for (int k = 1; k <= 2; ++k) {
            if ((k == 1 && !has_a) || (k == 2 && !has_b)) continue;
            // stage only operands that walk the input axis (a broadcast along it reads one
            // element per tile column, which phase 2 fetches directly)
            int64_t si = std::llabs(d.st[best_di][k]), sc = std::llabs(d.st[0][k]);
            if (si > 0 && si < sc) ch.stage |= (k == 1 ? 1 : 2);
        }
        return ch;
    }
    if (d.shape[0] >= (uint64_t)(EW_BLOCK * EW_VEC)) ch.kind = EwKind::Rows;
    return ch;
}

// Outputs at least this large are written with streaming stores: they cannot
// stay in the 256 MB last-level cache for a following kernel anyway.
constexpr uint64_t EW_NT_STORE_MIN = 128ull << 20;
// Tiled launches moving at least this many bytes have their tile order tuned.
constexpr uint64_t EW_TUNE_MIN = 32ull << 20;

// Call f(std::integral_constant<int, v>) for the runtime value v in [0, 8),
// instantiating only masks whose load bits are operands the body has.
template <int OPS, class F>
static void with_ntm(int v, F&& f) {
    auto one = [&](auto c) {
        constexpr int V = decltype(c)::value;
        if constexpr (((V & 3) & ~OPS) == 0) f(c);
    };
    switch (v & 7) {
        case 0: one(std::integral_constant<int, 0>{}); break;
        case 1: one(std::integral_constant<int, 1>{}); break;
        case 2: one(std::integral_constant<int, 2>{}); break;
        case 3: one(std::integral_constant<int, 3>{}); break;
        case 4: one(std::integral_constant<int, 4>{}); break;
        case 5: one(std::integral_constant<int, 5>{}); break;
        case 6: one(std::integral_constant<int, 6>{}); break;
        default: one(std::integral_constant<int, 7>{}); break;
    }
}

// Launch `body` over the plan, splitting the outermost axis so that every
// launch indexes fewer than 2^31 elements with 32-bit arithmetic.
// TILED = false: the LDS-tiled kernel is not instantiated for this body, and a
// tiled choice runs the generic kernel instead.
// Streaming hints (compile-time kernel variants): loads of operands read once,
// stores of outputs of at least EW_NT_STORE_MIN bytes. Bodies with NT_JOINT
// take both hints or neither (fewer variants).
// NTV = false: only the plain (hint-free) variant of the row kernel is instantiated, for bodies
// off the hot path.
template <class Body, bool TILED = true, bool NTV = true>
static void launch_ew(rstsr_hip_ctx* ctx, const Body& body, Dims d, EwChoice ch, bool can_time = false) {
    constexpr int OPS = Body::OPS;
    if (!TILED && ch.kind == EwKind::Tiled) ch.kind = EwKind::Generic;
    if (OPS == 0 && ch.kind == EwKind::Tiled) ch.kind = EwKind::Generic;
    if (d.nd < 0) return;
    if (d.nd == 0) {
        d.nd = 1;
        d.shape.push_back(1);
        d.st.push_back({0, 0, 0});
    }
    const uint64_t LIM = 1ull << 31;
    // Kernels index one launch with 32-bit arithmetic: an axis of 2^31 or more elements, or
    // inner axes (all but the outermost) holding 2^30 or more, are cut into pieces of 2^30
    // along that axis, each launched on its own with the operands' offsets advanced.
    {
        int cut = -1;
        for (int i = 0; i < d.nd; ++i)
            if (d.shape[i] >= LIM && (cut < 0 || d.shape[i] > d.shape[cut])) cut = i;
        if (cut < 0 && d.nd > 1) {
            uint64_t inner = 1;
            for (int i = 0; i + 1 < d.nd; ++i) inner *= d.shape[i];
            if (inner >= LIM / 2) {
                for (int i = 0; i + 1 < d.nd; ++i)
                    if (cut < 0 || d.shape[i] > d.shape[cut]) cut = i;
            }
        }
        if (cut >= 0) {
            const uint64_t n = d.shape[cut];
            uint64_t rest = 1;
            for (int i = 0; i < d.nd; ++i)
                if (i != cut && i + 1 < d.nd) rest *= d.shape[i];
            // pieces short enough that the inner product stays under 2^30
            const uint64_t step = std::max<uint64_t>(1, std::min<uint64_t>(1ull << 30, (LIM / 2) / std::max<uint64_t>(rest, 1)));
            for (uint64_t s0 = 0; s0 < n; s0 += step) {
                Dims dd = d;
                dd.shape[cut] = std::min(step, n - s0);
                for (int k = 0; k < 3; ++k) dd.off[k] += (int64_t)s0 * d.st[cut][k];
                launch_ew<Body, TILED, NTV>(ctx, body, dd, ch, false);
            }
            return;
        }
    }
    uint64_t all = 1;
    for (auto n : d.shape) all *= n;
    bool once[3] = {true, true, true};
    for (int k = 1; k < 3; ++k)
        for (int i = 0; i < d.nd; ++i)
            if (d.shape[i] > 1 && d.st[i][k] == 0) once[k] = false;
    const bool big = Body::OUT_SIZE > 0 && all * Body::OUT_SIZE >= EW_NT_STORE_MIN;
    int lmask = ((OPS & 1) && once[1] ? 1 : 0) | ((OPS & 2) && once[2] ? 2 : 0);
    int rows_ntm = lmask | (big ? 4 : 0);
    if (Body::NT_JOINT) rows_ntm = (lmask == OPS && big) ? (OPS | 4) : 0;

    uint64_t inner = 1;
    for (int i = 0; i + 1 < d.nd; ++i) inner *= d.shape[i];
    int last = d.nd - 1;
    uint64_t nlast = d.shape[last];
    uint64_t per = std::max<uint64_t>(1, (LIM / 2) / std::max<uint64_t>(inner, 1));
    if (ch.kind == EwKind::Tiled && ch.di == last) per = nlast;  // tile axis itself is small enough
    for (uint64_t s0 = 0; s0 < nlast; s0 += per) {
        uint64_t cnt = std::min(per, nlast - s0);
        Dims dd = d;
        dd.shape[last] = cnt;
        for (int k = 0; k < 3; ++k) dd.off[k] += (int64_t)s0 * d.st[last][k];
        EwPlan p = to_plan(dd);
        uint64_t total = inner * cnt;
        if (ch.kind == EwKind::Generic) {
            uint32_t nb = ceil_div_u32(total, EW_BLOCK * EW_VEC);
            hipLaunchKernelGGL(k_ew_generic<Body>, dim3(nb), dim3(EW_BLOCK), 0, ctx->stream, body, p, (uint32_t)total);
        } else if (ch.kind == EwKind::Rows) {
            uint32_t nchunk = ceil_div_u32(dd.shape[0], EW_BLOCK * EW_VEC);
            uint64_t nrows = total / dd.shape[0];
            uint64_t nb = nrows * nchunk;
            if (nb >= LIM) RSTSR_RAISE(UnImplemented, "grid too large");
            auto go = [&](auto ntm) {
                hipLaunchKernelGGL((k_ew_rows<Body, decltype(ntm)::value>), dim3((uint32_t)nb), dim3(EW_BLOCK), 0,
                                   ctx->stream, body, p, nchunk, FastDiv(nchunk));
            };
            if constexpr (!NTV) {
                go(std::integral_constant<int, 0>{});
            } else if constexpr (Body::NT_JOINT) {
                if (rows_ntm) go(std::integral_constant<int, OPS | 4>{});
                else go(std::integral_constant<int, 0>{});
            } else {
                with_ntm<OPS>(rows_ntm, go);
            }
        } else if constexpr (TILED && OPS != 0) {
            // tiles: 64x64 for 8-byte elements, 128 (input axis) x 64 for 4-byte and smaller, 32x32 for 16-byte
            constexpr size_t ES = sizeof(typename Body::TA) > sizeof(typename Body::TB) ? sizeof(typename Body::TA)
                                                                                          : sizeof(typename Body::TB);
            constexpr int TI = ES > 8 ? 32 : (ES > 4 ? 64 : 128);
            constexpr int TC = ES > 8 ? 32 : 64;
            uint32_t ntc = ceil_div_u32(dd.shape[0], TC);
            uint32_t nti = ceil_div_u32(dd.shape[ch.di], TI);
            uint64_t nouter = total / (dd.shape[0] * dd.shape[ch.di]);
            uint64_t nb = nouter * nti * ntc;
            if (nb >= LIM) RSTSR_RAISE(UnImplemented, "grid too large");
            int od = -1;
            for (int k = 1; k < dd.nd; ++k)
                if (k != ch.di) od = k;

            int xmap = 0;
            auto go = [&](auto stage, auto one, auto ntm) {
                hipLaunchKernelGGL((k_ew_tiled<Body, TI, TC, decltype(stage)::value, decltype(one)::value,
                                               decltype(ntm)::value>),
                                   dim3((uint32_t)nb), dim3(EW_BLOCK), 0, ctx->stream, body, p, ch.di, od, nti, ntc,
                                   FastDiv(nti), FastDiv(ntc), xmap);
            };
            // ONE (at most three axes): streaming variants, staged loads all-or-none; more axes: plain
            auto by_nt = [&](auto stage) {
                constexpr int S = decltype(stage)::value;
                if (dd.nd <= 3) {
                    const bool lo = ((S & 1) == 0 || once[1]) && ((S & 2) == 0 || once[2]);
                    using Y = std::true_type;
                    switch ((lo ? 1 : 0) | (big ? 2 : 0)) {
                        case 0: go(stage, Y{}, std::integral_constant<int, 0>{}); break;
                        case 1: go(stage, Y{}, std::integral_constant<int, S>{}); break;
                        case 2: go(stage, Y{}, std::integral_constant<int, 4>{}); break;
                        default: go(stage, Y{}, std::integral_constant<int, S | 4>{}); break;
                    }
                } else {
                    go(stage, std::false_type{}, std::integral_constant<int, 0>{});
                }
            };
            // choose() tiles along an input's unit axis, so that input is always staged:
            // a body with one array operand stages exactly that one
            auto launch = [&]() {
                if constexpr (OPS == 1) {
                    by_nt(std::integral_constant<int, 1>{});
                } else if constexpr (OPS == 2) {
                    by_nt(std::integral_constant<int, 2>{});
                } else {
                    switch (ch.stage) {
                        case 1: by_nt(std::integral_constant<int, 1>{}); break;
                        case 2: by_nt(std::integral_constant<int, 2>{}); break;
                        default: by_nt(std::integral_constant<int, 3>{}); break;
                    }
                }
            };
            // The fastest tile order depends on how the shapes and strides meet the
            // HBM channel interleave (measured: either order can win by 5-10 %), so it
            // is tuned per problem class on launches of at least EW_TUNE_MIN bytes.
            const uint64_t bytes = total * (sizeof(typename Body::TA) + Body::OUT_SIZE);
            if (can_time && per >= nlast && bytes >= EW_TUNE_MIN) {
                std::string key = "ew|" + std::to_string(sizeof(typename Body::TA)) + "|" +
                                  std::to_string(Body::OUT_SIZE) + "|" + std::to_string(ch.stage) + "|" +
                                  std::to_string(ntc) + "|" + std::to_string(nti) + "|" + std::to_string(nouter) +
                                  "|" + std::to_string(dd.st[0][1] + dd.st[0][2]) + "|" +
                                  std::to_string(dd.st[ch.di][0]);
                tune_select(ctx, key, 3, 0, true, [&](int c) {
                    xmap = c;
                    launch();
                    check_launch("elementwise");
                });
            } else {
                launch();
            }
        }
        check_launch("elementwise");
    }
}

/* =============================== dispatch =============================== */

template <class F>
static void dispatch_dtype(rstsr_dtype_t t, F&& f) {
    switch (t) {
        case RSTSR_F32: f(float{}); break;
        case RSTSR_F64: f(double{}); break;
        case RSTSR_I32: f(int32_t{}); break;
        case RSTSR_I64: f(int64_t{}); break;
        case RSTSR_C64: f(c64{}); break;
        case RSTSR_C128: f(c128{}); break;
        default: RSTSR_RAISE(UnImplemented, "dtype not supported for this operation");
    }
}

}  // namespace hip
}  // namespace rstsr
