// porosity.hip -- the coupling-scale porosity of Kernel_prep_sphdem
// (kernel/KERNEL_SPHDEM_interaction.py:74-96, :111) for uniform DEM h.
//
// For each fluid subject i the reference sums V_j W(|r_ij|, h_dem) over all
// fluid j within 2 h_dem (~2.7e5 particles at dx = 0.004), then divides each
// nearby DEM's V_b W by it.  Only subjects with a DEM within 2 h_b need the
// sum (elsewhere solid = 0 and the porosity is clamp(1)), so:
//   1. mark:   near-DEM flag per fluid slot; far subjects get clamp(1) here.
//   2. stream: one wave per 64 consecutive fluid slots holding a flagged
//              subject.  The wave walks the buckets of the union of its
//              subjects' balls in the reference's query order (z, then y,
//              then x, then sorted order inside a bucket), 64 candidates at a
//              time: one coalesced load per lane, staged in the wave's LDS
//              slot and read back by every lane.  Each lane adds the term
//              only if the candidate lies within 2 h_dem of its own subject.
//              Every candidate that passes the reference's distance test lies
//              in the union and inside the subject's own query box, so each
//              sum receives exactly the reference's terms in the reference's
//              order; buckets that cannot hold such a term are skipped.
// Math: EXACT evaluates W with IEEE division and square root as written in
// kernel/Kernel_KNL.py, in that order (bit-exact).  FAST uses the hardware
// square root, fused multiply-adds with 1/h folded in, the factored
// polynomial C u^4 (1+2q) with C applied once per sum, and u = 1 - q/2
// clamped to [0, 1] in place of the support test; in FP32 the squared
// distances come from the matrix cores and candidates that cannot reach any
// subject of the group are dropped before them (mfma_blocks and
// queue_range_mfma below), which changes the rounding of the sums, not
// their terms.
// FAST, FP32, symmetric (when the work struct brings fl_sum): W(r_ij) is the
// same for i's sum and j's, so a pair of flagged subjects in different waves
// is evaluated once, by the lower wave: it adds V_j W to i's sum and V_i W to
// j's, and the higher wave drops i from its stream.  The mirrored terms of a
// candidate are summed over the wave's subjects and added to j's fixed-point
// accumulator with a 64-bit integer atomic (order-free, so the result does
// not depend on which wave adds first); a last pass adds the subject's own
// sum and finishes the porosity.  Same terms, other rounding.
#include "common.hpp"

#include <climits>
#include <type_traits>
#include <cmath>
#include <limits>

namespace sphdem {


__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform_f(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T> __device__ __forceinline__ T wave_min(T v) {
    for (int off = 32; off > 0; off >>= 1) {
        const T o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}
template <typename T> __device__ __forceinline__ T wave_max(T v) {
    for (int off = 32; off > 0; off >>= 1) {
        const T o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Slot ranges of virtual x cells [xa, xb] (xa + 2^20 >= 0, at most dx cells)
// in a bucket row: two ranges when the row wraps around the folded grid.
__device__ __forceinline__ int row_slots(const int *lo, const int *hi, int base, int dx, int xa,
                                         int xb, int &a1, int &b1, int &a2, int &b2) {
    a1 = b1 = a2 = b2 = 0;
    if (xb < xa) return 0;
    const int len = xb - xa + 1;
    const int cx = fold(xa, dx);
    const int first = min(len, dx - cx);
    a1 = lo[base + cx];
    b1 = hi[base + cx + first - 1];
    if (len > first) {
        a2 = lo[base];
        b2 = hi[base + len - first - 1];
        return 2;
    }
    return 1;
}

__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }
__device__ __forceinline__ double clamp01(double x) { return fmin(fmax(x, 0.0), 1.0); }

// One term V W(|me - q|, h) of a neighbour sum.  FAST: u = 1 - q/2 clamped
// to [0, 1] stands in for the support test (W = 0 from q = 2 on), the 1/h
// scale folds into the two multiply-adds of u and 1 + 2q, and C is applied
// once per sum; a subject placed far away (1e15) adds nothing.
template <typename T, bool FAST>
__device__ __forceinline__ void add_term(T &acc, const V4<T> &me, T qx, T qy, T qz, T V, T thr2,
                                         T h, T inv_h) {
    const T dx = me.x - qx, dy = me.y - qy, dz = me.z - qz;
    if (FAST) {
        const T td = fma(dz, dz, fma(dy, dy, dx * dx));
        const T r = fast_sqrt(td);
        const T u = clamp01(fma((T)-0.5 * inv_h, r, (T)1.0));
        const T pq = fma((T)2.0 * inv_h, r, (T)1.0);
        T t = u * u;
        t = t * t;
        acc = fma(V, t * pq, acc);
    } else {
        const T td = dx * dx + dy * dy + dz * dz;
        if (td < thr2) acc = acc + V * wendland_w(msqrt(td), h);
    }
}

// Add the terms of sorted slots [a, b) for every lane of the wave.  The
// wave loads 64 candidates at a time with one coalesced load per lane
// (positions and m/rho packed as (x, y, z, V)), stages them in its own LDS
// slot and every lane reads them back as broadcasts, in slot order; the
// next 64 are loaded while these are processed.  Padding past b sits far
// away and fails the distance test.
template <typename T, bool FAST>
__device__ __forceinline__ void stream_range(T &acc, const V4<T> &me, int a, int b,
                                             const V4<T> *pos, const T *Vs, V4<T> *buf, int lane,
                                             T thr2, T h, T inv_h) {
    if (b <= a) return;
    // padding: far enough to fail the EXACT test, finite in FAST's terms
    const T far = FAST ? (T)1.0e15 : (T)1.0e30;
    auto load = [&](int s) {
        V4<T> c = {far, far, far, (T)0};
        if (s < b) {
            c = pos[s];
            c.w = Vs[s];
        }
        return c;
    };
    V4<T> cur = load(a + lane);
    for (int s0 = a; s0 < b; s0 += 64) {
        buf[lane] = cur;
        __builtin_amdgcn_wave_barrier();
        const V4<T> nxt = load(s0 + 64 + lane);
        const int cnt = uniform_i(min(64, b - s0));
        if (cnt == 64) {
#pragma unroll 16
            for (int j = 0; j < 64; ++j) {
                const V4<T> q = buf[j];
                add_term<T, FAST>(acc, me, q.x, q.y, q.z, q.w, thr2, h, inv_h);
            }
        } else {
#pragma unroll 4
            for (int j = 0; j < cnt; ++j) {
                const V4<T> q = buf[j];
                add_term<T, FAST>(acc, me, q.x, q.y, q.z, q.w, thr2, h, inv_h);
            }
        }
        __builtin_amdgcn_wave_barrier();
        cur = nxt;
    }
}

// FAST stream of sorted slots [a, b): candidates are staged relative to the
// group centre c as (x', y', z', |q'|^2) with V beside them, so that a term
// needs |r|^2 = (|m|^2 + |q'|^2) - 2 m.q' (one add, three multiply-adds,
// with mx2 = -2 m) instead of three differences and a dot product.  With
// |m|, |q'| <= ~0.2 m the rounding of that form stays ~1e-8 absolute, and
// |.| under the square root keeps a cancelled value finite.  Lanes outside
// the group and padding slots sit 1e15 away and add 0.
template <typename T>
__device__ __forceinline__ void stream_range_rel(T &acc, const V3<T> &mx2, T m2, int a, int b,
                                                 const V4<T> *pos, const T *Vs, V4<T> *buf,
                                                 T *vbuf, int lane, const V3<T> &c, T k1,
                                                 T k2) {
    if (b <= a) return;
    const T far = (T)1.0e15;
    auto load = [&](int s, T &v) {
        V4<T> q = {far, far, far, (T)0};
        v = (T)0;
        if (s < b) {
            q = pos[s];
            v = Vs[s];
        }
        q.x -= c.x;
        q.y -= c.y;
        q.z -= c.z;
        q.w = fma(q.z, q.z, fma(q.y, q.y, q.x * q.x));
        return q;
    };
    T vcur;
    V4<T> cur = load(a + lane, vcur);
    for (int s0 = a; s0 < b; s0 += 64) {
        buf[lane] = cur;
        vbuf[lane] = vcur;
        __builtin_amdgcn_wave_barrier();
        T vnxt;
        const V4<T> nxt = load(s0 + 64 + lane, vnxt);
        const int cnt = uniform_i(min(64, b - s0));
        auto term = [&](int j) {
            const V4<T> q = buf[j];
            const T td = fma(mx2.x, q.x, fma(mx2.y, q.y, fma(mx2.z, q.z, q.w + m2)));
            const T r = fast_sqrt(fabs(td));
            const T u = clamp01(fma(k1, r, (T)1.0));
            const T pq = fma(k2, r, (T)1.0);
            T t = u * u;
            t = t * t;
            acc = fma(vbuf[j], t * pq, acc);
        };
        if (cnt == 64) {
#pragma unroll 8
            for (int j = 0; j < 64; ++j) term(j);
        } else {
#pragma unroll 4
            for (int j = 0; j < cnt; ++j) term(j);
        }
        __builtin_amdgcn_wave_barrier();
        cur = nxt;
        vcur = vnxt;
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// FAST, FP32: the squared distances of a wave's 64 subjects to 16 staged
// candidates are four 16x16x4 FP32 matrix products (v_mfma_f32_16x16x4f32:
// lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds
// D[4(l/16)+t][l%16]).  Per tile T: A row i = (-2 m_s, 1) of subject
// s = 16T + i, B column j = (q'_j, |q'_j|^2), C = |m_s|^2, so
// D = |m|^2 + |q'|^2 - 2 m.q', the form of stream_range_rel.  The kernel
// term stays on the VALU: each lane adds its 16 (subject, candidate) terms
// into acc[T][t] (subject 16T + 4(l/16) + t, candidates l%16 mod 16), and the
// 16 partial sums of a subject are added across lanes once per group.  Here:
// nsub staged 16-candidate blocks, buf/vbuf holding candidates relative to the
// group centre as (x', y', z', |q'|^2) and V.
// The C operand of tile T: |m|^2 of subjects 16T + 4(l/16) + 0..3, read from
// the wave's LDS slot (one broadcast read per tile) instead of held in VGPRs.
struct COp {
    const float *m2;
    int comp;
    __device__ __forceinline__ f32x4 operator[](int T) const {
        return *reinterpret_cast<const f32x4 *>(m2 + 16 * T + 4 * comp);
    }
};

__device__ __forceinline__ void mfma_blocks(f32x4 (&acc)[4], const float (&a)[4],
                                            const COp &c, const V4<float> *buf,
                                            const float *vbuf, int nsub, int lane, float k1,
                                            float k2) {
    const int col = lane & 15, comp = lane >> 4;
    for (int sc = 0; sc < nsub; ++sc) {
        // component comp of candidate 16 sc + col, read at its address (a
        // four-way select here compiles to a per-lane branch tree)
        const float bv = reinterpret_cast<const float *>(buf + 16 * sc + col)[comp];
        const float vj = vbuf[16 * sc + col];
        // V (1 + 2q) = fma(2V/h, r, V): the lane's candidate V folded in once
        const float kv = k2 * vj;
        // all four products in flight before the first result is read; rows
        // of subjects outside the group sit 1e15 away and add 0
        f32x4 d[4];
#pragma unroll
        for (int T = 0; T < 4; ++T)
            d[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[T], bv, c[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < 4; ++T) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float r = fast_sqrt(fabsf(d[T][t]));
                const float u = clamp01(fmaf(k1, r, 1.0f));
                const float pqv = fmaf(kv, r, vj);
                float w = u * u;
                w = w * w;
                acc[T][t] = fmaf(w, pqv, acc[T][t]);
            }
        }
    }
}

// ---- symmetric FP32 fast path
// The flags, each flagged slot's work item (tag: the chunk list's, or NULL:
// slot / 64), the fixed-point accumulators, their scale (units per unit of
// the unscaled sum; a power of two) and the wave's own item.  Pair rule of a
// flagged candidate j: item(j) < cid: that item's pair (skipped here);
// item(j) > cid: evaluated here for both sums; item(j) == cid: each subject
// sums its own terms.
struct SymCtx {
    const int *flag;
    const int *tag;
    unsigned long long *col;
    float scale;
    int cid;
};

// -1 for an unflagged slot
__device__ __forceinline__ int item_of(const SymCtx &sym, int t) {
    return sym.tag ? sym.tag[t] : (sym.flag[t] != 0 ? t >> 6 : -1);
}

// v >= 0 in fixed point: x = v scale (exact), split at 2^32 exactly (the
// high part truncated, the low part an exact float difference by Sterbenz).
__device__ __forceinline__ unsigned long long to_fixed(float v, float scale) {
    const float x = v * scale;
    const float hi = truncf(x * 2.3283064365386963e-10f);
    const unsigned lo = (unsigned)(x - hi * 4294967296.0f);
    return ((unsigned long long)(unsigned)hi << 32) | lo;
}

// The sum of v over lanes l%16 + {0, 16, 32, 48}, in every lane: the
// halves' sums by a 32-lane swap, then the rows' by a 16-lane swap
// (v_permlane32_swap, v_permlane16_swap: no LDS round trip).
__device__ __forceinline__ float quarter_sum(float v) {
    const auto h = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v),
                                                    false, false);
    const float s = __uint_as_float(h[0]) + __uint_as_float(h[1]);
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(s), __float_as_uint(s),
                                                    false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// The symmetric queue's blocks: as mfma_blocks, and each lane also sums
// V_i W over its 16 (subject, candidate) pairs for its candidate (column
// l%16); the four quarters' sums are added and lanes 0..15 add the total to
// the candidate's accumulator.  vsub: V of the wave's subjects by lane (0
// outside the group); sid: the candidates' sorted slots (-1 for padding).
// (Issuing the next block's products before this block's terms, with one
// set of results or two, measured slower.)
__device__ __forceinline__ void mfma_blocks_sym(f32x4 (&acc)[4], const float (&a)[4],
                                                const COp &c, const V4<float> *buf,
                                                const float *vbuf, const int *sid, int nsub,
                                                int lane, float k1, float k2, const float *vsub,
                                                const SymCtx &sym) {
    const int col = lane & 15, comp = lane >> 4;
    for (int sc = 0; sc < nsub; ++sc) {
        const float bv = reinterpret_cast<const float *>(buf + 16 * sc + col)[comp];
        const float vj = vbuf[16 * sc + col];
        f32x4 d[4];
#pragma unroll
        for (int T = 0; T < 4; ++T)
            d[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[T], bv, c[T], 0, 0, 0);
        float cs = 0.0f;
#pragma unroll
        for (int T = 0; T < 4; ++T) {
            // V of subjects 16T + 4(l/16) + 0..3: one broadcast read per quarter
            const f32x4 vv = *reinterpret_cast<const f32x4 *>(vsub + 16 * T + 4 * comp);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float r = fast_sqrt(fabsf(d[T][t]));
                const float u = clamp01(fmaf(k1, r, 1.0f));
                const float pq = fmaf(k2, r, 1.0f);
                float w = u * u;
                w = w * w;
                w = w * pq;
                acc[T][t] = fmaf(vj, w, acc[T][t]);
                cs = fmaf(vv[t], w, cs);
            }
        }
        cs = quarter_sum(cs);
        if (lane < 16) {
            const int j = sid[16 * sc + col];
            if (j >= 0 && cs > 0.0f) atomicAdd(sym.col + j, to_fixed(cs, sym.scale));
        }
    }
}

// Symmetric queue of a wave (SYM): candidates, V, sorted slots and fill.
struct SymQueue {
    V4<float> *buf;
    float *vbuf;
    int *sid;
    float *vsub;   // [64] V of the wave's subjects by lane (0 outside the group)
    int fill;
    float *msub;   // [64] |m|^2 of the wave's subjects by lane (the products' C operand)
};

// FAST, FP32, culled: a candidate farther than the support from the box of
// the group's subjects adds 0 to every subject's sum, so it is dropped as it
// is loaded; the others queue in the wave's LDS slot (128 entries, `fill`
// of them pending, wave-uniform) and go through the matrix products 64 at a
// time.  The box is centred on gc with half-extents he, so its distance is
// measured on the relative coordinates the queue stores.  The test uses
// r2 = support^2 (1 + 1e-4) against a distance rounded to ~1e-7 m, so no
// candidate within the support is dropped.
template <bool SYM>
__device__ __forceinline__ void queue_range_mfma(f32x4 (&acc)[4], const float (&a)[4],
                                                 const COp &c, int &fill, int a0, int b0,
                                                 const V4<float> *pos, const float *Vs,
                                                 V4<float> *buf, float *vbuf, int lane,
                                                 const V3<float> &gc, const V3<float> &he,
                                                 float r2, float k1, float k2, SymQueue &sq,
                                                 const SymCtx &sym) {
    if (b0 <= a0) return;
    // slots past the range read its last slot again and are never kept
    auto load = [&](int s, float &v, int &f) {
        const int t = min(s, b0 - 1);
        v = Vs[t];
        f = SYM ? item_of(sym, t) : -1;
        return pos[t];
    };
    float vcur;
    int fcur;
    V4<float> cur = load(a0 + lane, vcur, fcur);
    for (int s0 = a0; s0 < b0; s0 += 64) {
        float vnxt;
        int fnxt;
        const V4<float> nxt = load(s0 + 64 + lane, vnxt, fnxt);
        V4<float> q;
        q.x = cur.x - gc.x;
        q.y = cur.y - gc.y;
        q.z = cur.z - gc.z;
        const float ex = fmaxf(fabsf(q.x) - he.x, 0.0f);
        const float ey = fmaxf(fabsf(q.y) - he.y, 0.0f);
        const float ez = fmaxf(fabsf(q.z) - he.z, 0.0f);
        const int s = s0 + lane;
        const bool near = s < b0 && fmaf(ez, ez, fmaf(ey, ey, ex * ex)) <= r2;
        // SYM: a flagged candidate of a lower item is that item's pair; one
        // of a higher item goes to the symmetric queue
        const bool higher = SYM && fcur > sym.cid;
        const bool keep = near && !(SYM && fcur >= 0 && fcur != sym.cid);
        q.w = fmaf(q.z, q.z, fmaf(q.y, q.y, q.x * q.x));
        const unsigned long long mask = __ballot(keep);
        if (keep) {
            const int at = fill + (int)__builtin_amdgcn_mbcnt_hi(
                                      (unsigned)(mask >> 32),
                                      __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
            buf[at] = q;
            vbuf[at] = vcur;
        }
        fill = uniform_i(fill + __popcll(mask));
        if constexpr (SYM) {
            const bool keep_s = near && higher;
            const unsigned long long ms = __ballot(keep_s);
            if (keep_s) {
                const int at = sq.fill + (int)__builtin_amdgcn_mbcnt_hi(
                                             (unsigned)(ms >> 32),
                                             __builtin_amdgcn_mbcnt_lo((unsigned)ms, 0u));
                sq.buf[at] = q;
                sq.vbuf[at] = vcur;
                sq.sid[at] = s;
            }
            sq.fill = uniform_i(sq.fill + __popcll(ms));
            if (sq.fill >= 64) {
                __builtin_amdgcn_wave_barrier();
                mfma_blocks_sym(acc, a, c, sq.buf, sq.vbuf, sq.sid, 4, lane, k1, k2, sq.vsub,
                                sym);
                __builtin_amdgcn_wave_barrier();
                if (lane < sq.fill - 64) {
                    sq.buf[lane] = sq.buf[64 + lane];
                    sq.vbuf[lane] = sq.vbuf[64 + lane];
                    sq.sid[lane] = sq.sid[64 + lane];
                }
                sq.fill -= 64;
                __builtin_amdgcn_wave_barrier();
            }
        }
        if (fill >= 64) {
            __builtin_amdgcn_wave_barrier();
            mfma_blocks(acc, a, c, buf, vbuf, 4, lane, k1, k2);
            __builtin_amdgcn_wave_barrier();
            // entries 64 .. fill-1 move down to 0 .. fill-65 (disjoint ranges)
            if (lane < fill - 64) {
                buf[lane] = buf[64 + lane];
                vbuf[lane] = vbuf[64 + lane];
            }
            fill -= 64;
            __builtin_amdgcn_wave_barrier();
        }
        cur = nxt;
        vcur = vnxt;
        fcur = fnxt;
    }
}

// The symmetric queue's last entries, padded to a multiple of 16 with
// candidates 1e15 away (V = 0, no slot).
__device__ __forceinline__ void flush_queue_sym(f32x4 (&acc)[4], const float (&a)[4],
                                                const COp &c, SymQueue &sq, int lane,
                                                float k1, float k2, const SymCtx &sym) {
    if (sq.fill == 0) return;
    const int nsub = (sq.fill + 15) >> 4;
    if (lane >= sq.fill && lane < 16 * nsub) {
        const float far = 1.0e15f;
        sq.buf[lane] = V4<float>{far, far, far, 3.0e30f};
        sq.vbuf[lane] = 0.0f;
        sq.sid[lane] = -1;
    }
    __builtin_amdgcn_wave_barrier();
    mfma_blocks_sym(acc, a, c, sq.buf, sq.vbuf, sq.sid, nsub, lane, k1, k2, sq.vsub, sym);
    __builtin_amdgcn_wave_barrier();
    sq.fill = 0;
}

// The queue's last entries, padded to a multiple of 16 with candidates
// 1e15 away (V = 0), which add 0.
__device__ __forceinline__ void flush_queue_mfma(f32x4 (&acc)[4], const float (&a)[4],
                                                 const COp &c, int &fill,
                                                 V4<float> *buf, float *vbuf, int lane,
                                                 float k1, float k2) {
    if (fill == 0) return;
    const int nsub = (fill + 15) >> 4;
    if (lane >= fill && lane < 16 * nsub) {
        const float far = 1.0e15f;
        buf[lane] = V4<float>{far, far, far, 3.0e30f};
        vbuf[lane] = 0.0f;
    }
    __builtin_amdgcn_wave_barrier();
    mfma_blocks(acc, a, c, buf, vbuf, nsub, lane, k1, k2);
    __builtin_amdgcn_wave_barrier();
    fill = 0;
}

// 1. near-DEM flag; far subjects get their porosity now.
template <typename T>
__global__ void k_por_mark(Prm<T> p, SphV<T> P, DemV<T> D, GridV<T> gs, GridV<T> gd, T csup,
                           int *flag, int *dem_near, int remap) {
    const int k = subject_id(remap);
    if (k >= gs.n) return;
    const V4<T> me = gs.s[k];
    gd = at_inst(gd, inst_of(gs, k));
    bool near = false;
    query_runs(gd, me.x, me.y, me.z, csup, [&](int lo, int hi) {
        for (int s = lo; s < hi; ++s) {
            const V4<T> q = gd.s[s];
            const T dist = length(mk3(me.x - q.x, me.y - q.y, me.z - q.z));
            if (dist < (T)2.0 * D.h[gd.ids[s]]) {
                near = true;
                // the DEM's own fluid search (k_interaction_dem) tests the
                // same distance of the same two stored positions
                if (dem_near) dem_near[gd.ids[s]] = 1;
            }
        }
    });
    flag[k] = near ? 1 : 0;
    if (!near) P.porosity[gs.ids[k]] = clampr((T)1.0, p.eps_min, p.eps_max);
}

// KERNEL_SPHDEM_interaction.py:85-96, 111: the solid fraction of the DEM
// within 2 h_b of `me` over the fluid sum flt, and the clamped porosity.
template <typename T>
__device__ __forceinline__ T porosity_of(const Prm<T> &p, const DemV<T> &D, const GridV<T> &gd,
                                         const V4<T> &me, T flt, T csup) {
    T solid = (T)0;
    query_runs(gd, me.x, me.y, me.z, csup, [&](int lo, int hi) {
        for (int s = lo; s < hi; ++s) {
            const int b = gd.ids[s];
            const V4<T> q = gd.s[s];
            const T hb = D.h[b];
            const T dist = length(mk3(me.x - q.x, me.y - q.y, me.z - q.z));
            if (dist < (T)2.0 * hb)
                solid = solid + D.volume[b] * wendland_w(dist, hb) / (flt + (T)1.0e-20);
        }
    });
    return clampr((T)1.0 - solid, p.eps_min, p.eps_max);
}

// Stream the union of the balls of the wave's lanes with `in` set; other
// lanes compare against a zero threshold and add nothing.  Returns false
// (having done nothing) if that union spans more than dims - 2 buckets of
// the folded grid along some axis.  FP32 FAST: only the bucket rows
// part, part + parts, ... of the union's query (parts > 1: a partial sum).
template <typename T, bool FAST, bool SYM>
__device__ bool stream_group(T &flt, const V4<T> &me, bool in, const GridV<T> &gs, const T *Vs,
                             V4<T> *wbuf, T *vbuf, int lane, T csup, T thr2, T h, T inv_h,
                             SymQueue &sq, T vme, const SymCtx &sym, int part, int parts) {
    const T big = std::numeric_limits<T>::max();
    const T bxl = uniform_f(wave_min(in ? me.x : big)), bxh = uniform_f(wave_max(in ? me.x : -big));
    const T byl = uniform_f(wave_min(in ? me.y : big)), byh = uniform_f(wave_max(in ? me.y : -big));
    const T bzl = uniform_f(wave_min(in ? me.z : big)), bzh = uniform_f(wave_max(in ? me.z : -big));
    const T inv = gs.inv, w = (T)1.0 / inv;
    const T r2c = csup * csup * (T)1.0001;
    // union of the lanes' query boxes: int() is monotone, so the box of the
    // extreme positions bounds every lane's [int((p-r)/w), int((p+r)/w)]
    const int X0 = uniform_i((int)((bxl - csup) * inv)), X1 = uniform_i((int)((bxh + csup) * inv));
    const int Y0 = uniform_i((int)((byl - csup) * inv)), Y1 = uniform_i((int)((byh + csup) * inv));
    const int Z0 = uniform_i((int)((bzl - csup) * inv)), Z1 = uniform_i((int)((bzh + csup) * inv));
    // at most dims - 2 cells per axis, so every other fold of a bucket lies
    // two cells outside the box and the culling below drops no term
    if (!(X1 - X0 + 3 <= gs.dx && Y1 - Y0 + 3 <= gs.dy && Z1 - Z0 + 3 <= gs.dz &&
          X0 + kOrigin >= 0 && Y0 + kOrigin >= 0 && Z0 + kOrigin >= 0))
        return false;
    const T thr = in ? thr2 : (T)0;
    // FAST: the subject relative to the group centre (far away if not in)
    const V3<T> gc = mk3(uniform_f((bxl + bxh) * (T)0.5), uniform_f((byl + byh) * (T)0.5),
                         uniform_f((bzl + bzh) * (T)0.5));
    const T far = (T)1.0e15;
    const V3<T> m = in ? mk3(me.x - gc.x, me.y - gc.y, me.z - gc.z) : mk3(far, far, far);
    const V3<T> mx2 = (T)-2.0 * m;
    const T m2 = dot(m, m);
    const T k1 = (T)-0.5 * inv_h, k2 = (T)2.0 * inv_h;
    const int dxy = gs.dx * gs.dy;
    // FP32 FAST: matrix-core distances (mfma_blocks, fed by queue_range_mfma); set up the
    // group's A rows and C values, accumulate, reduce at the end
    constexpr bool kMfma = FAST && std::is_same<T, float>::value;
    static_assert(!SYM || kMfma, "the symmetric sums are an FP32 fast-path form");
    const int col = lane & 15, comp = lane >> 4;
    float ma[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const COp mc{sq.msub, comp};
    f32x4 acc[4] = {};
    int fill = 0;
    if constexpr (SYM) {
        sq.vsub[lane] = in ? (float)vme : 0.0f;
        __builtin_amdgcn_wave_barrier();
    }
    if constexpr (kMfma) {
#pragma unroll
        for (int Tt = 0; Tt < 4; ++Tt) {
            const int sl = 16 * Tt + col;
            const float mxs = __shfl((float)m.x, sl, 64), mys = __shfl((float)m.y, sl, 64),
                        mzs = __shfl((float)m.z, sl, 64);
            ma[Tt] = comp == 0 ? -2.0f * mxs : comp == 1 ? -2.0f * mys
                                               : comp == 2 ? -2.0f * mzs : 1.0f;
        }
        sq.msub[lane] = (float)m2;
        __builtin_amdgcn_wave_barrier();
    }
    if constexpr (kMfma) {
        // FP32 FAST: the lanes find the slot ranges of 64 bucket rows at once
        // (row r0 + lane, z-major as the loops below), so the bucket-bound
        // loads of a row are not waited for one row at a time; the wave then
        // queues the rows' candidates in order
        const V3<float> gcf = mk3((float)gc.x, (float)gc.y, (float)gc.z);
        const V3<float> hef = mk3((float)((bxh - bxl) * (T)0.5), (float)((byh - byl) * (T)0.5),
                                  (float)((bzh - bzl) * (T)0.5));
        const int ny = Y1 - Y0 + 1, nrows = (Z1 - Z0 + 1) * ny;
        const int mine = part < nrows ? (nrows - part + parts - 1) / parts : 0;
        for (int r0 = 0; r0 < mine; r0 += 64) {
            int ra1 = 0, rb1 = 0, ra2 = 0, rb2 = 0;
            const int idx = (r0 + lane) * parts + part;
            if (r0 + lane < mine) {
                const int z = Z0 + idx / ny, y = Y0 + idx % ny;
                const T gz = axis_gap(bzl, bzh, z, w);
                const T gy = axis_gap(byl, byh, y, w);
                const T dd = gy * gy + gz * gz;
                if (dd <= r2c) {
                    const T chord = msqrt(r2c - dd);
                    const int xa = max(X0, (int)((bxl - chord) * inv) - 1);
                    const int xb = min(X1, (int)((bxh + chord) * inv) + 1);
                    const int row = fold(z, gs.dz) * dxy + fold(y, gs.dy) * gs.dx;
                    row_slots(gs.lo + gs.off, gs.hi + gs.off, row, gs.dx, xa, xb, ra1, rb1, ra2,
                              rb2);
                }
            }
            const int nr = min(64, mine - r0);
            for (int r = 0; r < nr; ++r) {
                const int a1 = __builtin_amdgcn_readlane(ra1, r);
                const int b1 = __builtin_amdgcn_readlane(rb1, r);
                const int a2 = __builtin_amdgcn_readlane(ra2, r);
                const int b2 = __builtin_amdgcn_readlane(rb2, r);
                queue_range_mfma<SYM>(acc, ma, mc, fill, a1, b1, (const V4<float> *)gs.s,
                                      (const float *)Vs, (V4<float> *)wbuf, (float *)vbuf, lane,
                                      gcf, hef, (float)r2c, (float)k1, (float)k2, sq, sym);
                queue_range_mfma<SYM>(acc, ma, mc, fill, a2, b2, (const V4<float> *)gs.s,
                                      (const float *)Vs, (V4<float> *)wbuf, (float *)vbuf, lane,
                                      gcf, hef, (float)r2c, (float)k1, (float)k2, sq, sym);
            }
        }
    } else {
        for (int z = Z0; z <= Z1; ++z) {
            const T gz = axis_gap(bzl, bzh, z, w);
            if (gz * gz > r2c) continue;
            const int cz = fold(z, gs.dz);
            for (int y = Y0; y <= Y1; ++y) {
                const T gy = axis_gap(byl, byh, y, w);
                const T dd = gy * gy + gz * gz;
                if (dd > r2c) continue;
                const T chord = msqrt(r2c - dd);
                const int xa = uniform_i(max(X0, (int)((bxl - chord) * inv) - 1));
                const int xb = uniform_i(min(X1, (int)((bxh + chord) * inv) + 1));
                if (xb < xa) continue;
                const int row = cz * dxy + fold(y, gs.dy) * gs.dx;
                int a1, b1, a2, b2;
                row_slots(gs.lo + gs.off, gs.hi + gs.off, row, gs.dx, xa, xb, a1, b1, a2, b2);
                a1 = uniform_i(a1); b1 = uniform_i(b1);
                a2 = uniform_i(a2); b2 = uniform_i(b2);
                if (FAST) {
                    stream_range_rel(flt, mx2, m2, a1, b1, gs.s, Vs, wbuf, vbuf, lane, gc, k1,
                                     k2);
                    stream_range_rel(flt, mx2, m2, a2, b2, gs.s, Vs, wbuf, vbuf, lane, gc, k1,
                                     k2);
                } else {
                    stream_range<T, FAST>(flt, me, a1, b1, gs.s, Vs, wbuf, lane, thr, h, inv_h);
                    stream_range<T, FAST>(flt, me, a2, b2, gs.s, Vs, wbuf, lane, thr, h, inv_h);
                }
            }
        }
    }
    if constexpr (kMfma) {
        flush_queue_mfma(acc, ma, mc, fill, (V4<float> *)wbuf, (float *)vbuf, lane, (float)k1,
                         (float)k2);
        if constexpr (SYM) flush_queue_sym(acc, ma, mc, sq, lane, (float)k1, (float)k2, sym);
        // subject 16T + 4(l/16) + t: add its 16 column partial sums, lanes
        // l%16 = 0..15 of each quarter (fixed xor tree), then hand the total
        // to the subject's own lane through the wave's LDS slot
#pragma unroll
        for (int Tt = 0; Tt < 4; ++Tt) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                float v = acc[Tt][t];
                v += __shfl_xor(v, 1, 64);
                v += __shfl_xor(v, 2, 64);
                v += __shfl_xor(v, 4, 64);
                v += __shfl_xor(v, 8, 64);
                acc[Tt][t] = v;
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (col == 0) {
#pragma unroll
            for (int Tt = 0; Tt < 4; ++Tt)
#pragma unroll
                for (int t = 0; t < 4; ++t) ((float *)vbuf)[16 * Tt + 4 * comp + t] = acc[Tt][t];
        }
        __builtin_amdgcn_wave_barrier();
        if (in) flt += (T)((const float *)vbuf)[lane];
        __builtin_amdgcn_wave_barrier();
    }
    return true;
}

// 2. the stream.  A wave takes one work item: 64 near-DEM subjects of one
// bucket row (the chunk list), or without the list 64 consecutive fluid
// slots.  The wave's subjects are streamed in groups that share an instance
// and a bucket row (y, z) -- 64 consecutive slots can straddle two rows --
// and, when a row group is still too wide (folded buckets), in sub-groups of
// 2^xt buckets along x; a sub-group that still does not fit (grids narrower
// than the query) walks its own query per lane.
// Lane subject: slot k where active; cid: the wave's work item (SYM's pair
// rule), tag: the chunk of each flagged slot or NULL (slot / 64).  SYM: the
// item's waves part = 0..parts-1 each stream their share of the bucket rows
// and add their partial sums to the fixed-point accumulators (the per-lane
// query of a group that fits no stream is part 0's).
template <typename T, bool FAST, bool SYM>
__device__ void por_wave(const Prm<T> &p, const SphV<T> &P, const DemV<T> &D, const GridV<T> &gs,
                         GridV<T> gd, const T *Vs, const int *flag, const int *tag, T csup, T hu,
                         T thr2, int xt, unsigned long long *col_sum, float col_scale, int k,
                         bool active, int cid, int part, int parts, int lane, V4<T> *wbuf,
                         T *vbuf, SymQueue sq) {
    const SymCtx sym{flag, tag, col_sum, col_scale, cid};
    const unsigned long long live = __ballot(active);
    if (live == 0ull) return;
    const int lead = __shfl(k, __ffsll((long long)live) - 1, 64);
    const int ks = active ? k : lead;
    const V4<T> me = gs.s[ks];
    const T vme = SYM ? Vs[ks] : (T)0;
    const int ik = inst_of(gs, ks);
    gd = at_inst(gd, ik);
    const T h = hu, inv_h = (T)1.0 / hu, inv = gs.inv;
    const int ky = (int)(me.y * inv), kz = (int)(me.z * inv), kx = ((int)(me.x * inv)) >> xt;
    T flt = (T)0;
    unsigned long long todo = live;
    while (todo) {
        const int ld = __ffsll((long long)todo) - 1;
        const int ly = __shfl(ky, ld, 64), lz = __shfl(kz, ld, 64), li = __shfl(ik, ld, 64);
        const unsigned long long grp =
            __ballot(active && ky == ly && kz == lz && ik == li) & todo;
        const bool in = (grp >> lane) & 1ull;
        const GridV<T> gg = at_inst(gs, uniform_i(li));
        if (!stream_group<T, FAST, SYM>(flt, me, in, gg, Vs, wbuf, vbuf, lane, csup, thr2, h,
                                        inv_h, sq, vme, sym, part, parts)) {
            unsigned long long rest = grp;
            while (rest) {
                const int l2 = __ffsll((long long)rest) - 1;
                const int lx = __shfl(kx, l2, 64);
                const unsigned long long sub = __ballot(kx == lx) & rest;
                const bool in2 = (sub >> lane) & 1ull;
                if (!stream_group<T, FAST, SYM>(flt, me, in2, gg, Vs, wbuf, vbuf, lane, csup,
                                                thr2, h, inv_h, sq, vme, sym, part, parts) &&
                    in2 && part == 0) {
                    query_runs(at_inst(gs, ik), me.x, me.y, me.z, csup, [&](int lo, int hi) {
                        for (int s = lo; s < hi; ++s) {
                            const V4<T> q = gs.s[s];
                            if constexpr (SYM) {
                                // the pair rule of the streamed groups, per pair
                                const int it = item_of(sym, s);
                                if (it >= 0 && it < sym.cid) continue;
                                if (it > sym.cid) {
                                    T w = (T)0;
                                    add_term<T, FAST>(w, me, q.x, q.y, q.z, (T)1, thr2, h,
                                                      inv_h);
                                    flt = fma(Vs[s], w, flt);
                                    const float cv = (float)(vme * w);
                                    if (cv > 0.0f)
                                        atomicAdd(sym.col + s, to_fixed(cv, sym.scale));
                                    continue;
                                }
                            }
                            add_term<T, FAST>(flt, me, q.x, q.y, q.z, Vs[s], thr2, h, inv_h);
                        }
                    });
                }
                rest &= ~sub;
            }
        }
        todo &= ~grp;
    }
    if (!active) return;
    if constexpr (SYM) {
        // the subject's own share; k_por_finish reads it with the mirrored terms
        if (flt > (T)0) atomicAdd(col_sum + k, to_fixed((float)flt, col_scale));
        return;
    }
    if (FAST) flt = flt * ((T)21.0 / ((T)16.0 * (T)SPHDEM_PI * hu * hu * hu));
    P.porosity[gs.ids[k]] = porosity_of(p, D, gd, me, flt, csup);
}

// The chunk list (CHUNKS): cl = {list, L, tag}, count = {chunks, flagged};
// wave c of the grid takes chunks c, c + waves, ...
struct ChunkV {
    const int *list, *L, *tag;
    int *count;
};

template <typename T, bool FAST, bool SYM, bool CHUNKS>
// At least 3 waves per SIMD: 168 VGPRs; the FP32 fast kernels then spill a
// few registers in their per-item and per-group setup, none in the
// candidate and product loops (measured at c2: 14.60 -> 12.54 ms against 2
// waves per SIMD; 4 waves, 128 VGPRs, spill in the loops: 14.73 ms).
__global__ __launch_bounds__(256, 3) void k_por_stream(Prm<T> p, SphV<T> P, DemV<T> D, GridV<T> gs,
                                                    GridV<T> gd, const T *Vs, const int *flag,
                                                    T csup, T hu, T thr2, int xt, int remap,
                                                    unsigned long long *col_sum,
                                                    float col_scale, ChunkV cl, int parts,
                                                    int queue) {
    // per wave: 64 staged candidates, 128 for the FP32 matrix-core queue;
    // SYM: a second 128-entry queue for the pairs this wave evaluates for both
    constexpr int kSlot = (FAST && std::is_same<T, float>::value) ? 128 : 64;
    constexpr int kSymSlot = SYM ? 128 : 1;
    __shared__ V4<T> stage[4 * kSlot];
    __shared__ T vstage[4 * kSlot];
    __shared__ V4<float> sstage[4 * kSymSlot];
    __shared__ float svstage[4 * kSymSlot];
    __shared__ int sidstage[4 * kSymSlot];
    __shared__ __align__(16) float vsubstage[4 * (SYM ? 64 : 1)];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    V4<T> *wbuf = stage + wv * kSlot;
    T *vbuf = vstage + wv * kSlot;
    __shared__ __align__(16) float msubstage[4 * 64];
    const SymQueue sq{sstage + wv * kSymSlot, svstage + wv * kSymSlot, sidstage + wv * kSymSlot,
                      vsubstage + wv * (SYM ? 64 : 1), 0, msubstage + wv * 64};
    const int wave = subject_id(remap) >> 6;
    if constexpr (CHUNKS) {
        // items (chunk, part) in list order, the parts of a chunk next to
        // each other.  queue 0: wave w takes items w, w + waves, ...;
        // queue 1: each wave takes the next item from the counter count[2]
        // (zero at the start of the pass) whenever it finishes one, so the
        // waves of a workgroup do not wait for each other and the pass ends
        // within about one item of the last wave's start.
        const int nch = cl.count[0], nfl = cl.count[1];
        const int total = nch * parts, waves = (gridDim.x * blockDim.x) >> 6;
        int item = wave;
        for (;;) {
            if (queue != 0) {
                int got = 0;
                if (lane == 0) got = atomicAdd(cl.count + 2, 1);
                item = __builtin_amdgcn_readfirstlane(__shfl(got, 0, 64));
            }
            if (item >= total) break;
            const int c = item / parts, part = item - c * parts;
            const int i = cl.list[c] + lane;
            const bool active = i < (c + 1 < nch ? cl.list[c + 1] : nfl);
            const int k = active ? cl.L[i] : 0;
            por_wave<T, FAST, SYM>(p, P, D, gs, gd, Vs, flag, cl.tag, csup, hu, thr2, xt,
                                   col_sum, col_scale, k, active, c, part, parts, lane, wbuf,
                                   vbuf, sq);
            if (queue == 0) item += waves;
        }
    } else {
        const int k = wave * 64 + lane;
        const bool active = k < gs.n && flag[k] != 0;
        por_wave<T, FAST, SYM>(p, P, D, gs, gd, Vs, flag, nullptr, csup, hu, thr2, xt, col_sum,
                               col_scale, k, active, wave, 0, 1, lane, wbuf, vbuf, sq);
    }
}

// ---- the chunk list: the near-DEM slots of each bucket row (instance,
// unfolded y and z cell: the groups the stream forms), compacted in slot
// order and cut every 64.  Arrays of n ints (fl_chunk, 7 of them):
//   F   exclusive scan of the flags; L[F[k]] = k: the flagged slots
//   R   exclusive scan of the row heads of L (R[i] + head(i) - 1: row of i)
//   RF  first index of L of each row
//   CS  exclusive scan of the chunk starts (i - RF[row]) % 64 == 0
//   list[chunk] = its first index of L;  tag[k] = chunk of flagged slot k
//   (-1 unflagged);  count[0] = chunks, count[1] = flagged slots,
//   count[2] = the stream's item counter (zeroed here).
template <typename T>
__device__ __forceinline__ bool row_head(const GridV<T> &gs, const int *L, int i) {
    if (i == 0) return true;
    const int k = L[i], j = L[i - 1];
    const V4<T> a = gs.s[k], b = gs.s[j];
    const T inv = gs.inv;
    return inst_of(gs, k) != inst_of(gs, j) || (int)(a.y * inv) != (int)(b.y * inv) ||
           (int)(a.z * inv) != (int)(b.z * inv);
}

__global__ void k_cl_compact(int n, const int *flag, const int *F, int *L, int *tag,
                             int *count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (k == 0) count[2] = 0;   // the stream's item counter
    if (flag[k] != 0) L[F[k]] = k;
    else tag[k] = -1;
    if (k == n - 1) count[1] = F[k] + (flag[k] != 0 ? 1 : 0);
}

template <typename T>
__global__ void k_cl_heads(GridV<T> gs, const int *L, const int *count, int *H) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < gs.n) H[i] = (i < count[1] && row_head(gs, L, i)) ? 1 : 0;
}

template <typename T>
__global__ void k_cl_firsts(GridV<T> gs, const int *L, const int *count, const int *R, int *RF) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count[1] && row_head(gs, L, i)) RF[R[i]] = i;
}

template <typename T>
__device__ __forceinline__ bool chunk_start(const GridV<T> &gs, const int *L, const int *R,
                                            const int *RF, int i) {
    const int row = R[i] + (row_head(gs, L, i) ? 1 : 0) - 1;
    return ((i - RF[row]) & 63) == 0;
}

template <typename T>
__global__ void k_cl_marks(GridV<T> gs, const int *L, const int *count, const int *R,
                           const int *RF, int *C) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < gs.n) C[i] = (i < count[1] && chunk_start(gs, L, R, RF, i)) ? 1 : 0;
}

template <typename T>
__global__ void k_cl_list(GridV<T> gs, const int *L, const int *R, const int *RF, const int *CS,
                          int *list, int *tag, int *count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nfl = count[1];
    if (i == 0 && nfl == 0) count[0] = 0;
    if (i >= nfl) return;
    const bool st = chunk_start(gs, L, R, RF, i);
    const int c = CS[i] + (st ? 1 : 0) - 1;
    if (st) list[c] = i;
    tag[L[i]] = c;
    if (i == nfl - 1) count[0] = c + 1;
}

// SYM: each flagged subject's sum is the fixed-point total of its own
// partial sums and the mirrored terms the lower items added (read, then
// cleared for the next step; the unscaled sum is left in row_sum), then the
// porosity as k_por_stream finishes it.
template <typename T>
__global__ void k_por_finish(Prm<T> p, SphV<T> P, DemV<T> D, GridV<T> gs, GridV<T> gd,
                             const int *flag, T *row_sum, unsigned long long *col_sum,
                             double inv_scale, T csup, T hu, int remap) {
    const int k = subject_id(remap);
    if (k >= gs.n || flag[k] == 0) return;
    const unsigned long long cv = col_sum[k];
    col_sum[k] = 0ull;
    T flt = (T)((double)cv * inv_scale);
    row_sum[k] = flt;
    flt = flt * ((T)21.0 / ((T)16.0 * (T)SPHDEM_PI * hu * hu * hu));
    const V4<T> me = gs.s[k];
    gd = at_inst(gd, inst_of(gs, k));
    P.porosity[gs.ids[k]] = porosity_of(p, D, gd, me, flt, csup);
}

// Smallest x with sqrt(x) >= s (IEEE, round to nearest): sqrt(td) < s is
// then exactly td < thr2.
template <typename T> T sqrt_threshold(T s) {
    const T big = std::numeric_limits<T>::max();
    T x = s * s;
    while (x > (T)0 && std::sqrt(std::nextafter(x, (T)0)) >= s) x = std::nextafter(x, (T)0);
    while (std::sqrt(x) < s) x = std::nextafter(x, big);
    return x;
}

template <typename T>
int launch_porosity_tiled(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_dem_t *D,
                          const sphdem_grid_t *gs, const sphdem_grid_t *gd,
                          const sphdem_work_t *w, double csup, double hu, int *dem_near,
                          hipStream_t st, hipEvent_t marked) {
    const int n = gs->n;
    if (n == 0) return SPHDEM_OK;
    ProfScope prof(SPHDEM_PASS_POROSITY, st);
    const T cs = (T)csup;
    const int bt = tuned_block(SPHDEM_PASS_POROSITY), rm = tuned_remap(SPHDEM_PASS_POROSITY);
    const int xt = tuning().por_x_tile;
    hipLaunchKernelGGL(k_por_mark<T>, dim3(blocks_for(n, bt)), dim3(bt), 0, st, prm<T>(p),
                       view<T>(P), view<T>(D), view<T>(gs), view<T>(gd), cs, w->fl_flag, dem_near,
                       rm);
    // the flags are final here: work that needs only them may start
    if (marked) (void)hipEventRecord(marked, st);
    const T thr2 = sqrt_threshold<T>(cs);
    // the chunk list of the flagged slots (fl_chunk: F, L, R, RF, CS, list,
    // tag of n ints each; the counts in fl_count[0..1])
    ChunkV cl{nullptr, nullptr, nullptr, nullptr};
    const bool chunked = w->fl_chunk != nullptr;
    if (chunked) {
        int *F = w->fl_chunk, *L = F + n, *R = L + n, *RF = R + n, *CS = RF + n, *list = CS + n,
            *tag = list + n;
        const dim3 g(blocks_for(n, 256)), b(256);
        int rc = exclusive_scan(w->fl_flag, F, n, w->fl_scan, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_cl_compact, g, b, 0, st, n, w->fl_flag, F, L, tag, w->fl_count);
        hipLaunchKernelGGL(k_cl_heads<T>, g, b, 0, st, view<T>(gs), L, w->fl_count, R);
        if ((rc = exclusive_scan(R, R, n, w->fl_scan, st))) return rc;
        hipLaunchKernelGGL(k_cl_firsts<T>, g, b, 0, st, view<T>(gs), L, w->fl_count, R, RF);
        hipLaunchKernelGGL(k_cl_marks<T>, g, b, 0, st, view<T>(gs), L, w->fl_count, R, RF, CS);
        if ((rc = exclusive_scan(CS, CS, n, w->fl_scan, st))) return rc;
        hipLaunchKernelGGL(k_cl_list<T>, g, b, 0, st, view<T>(gs), L, R, RF, CS, list, tag,
                           w->fl_count);
        cl = ChunkV{list, L, tag, w->fl_count};
    }
    const bool sym = p->fast_math && std::is_same<T, float>::value && w->fl_sum;
    if (sym) {
        // fixed-point units: the unscaled sum of a full ball (1/C) near 2^40
        const int e = 40 - (int)std::ceil(std::log2(16.0 * SPHDEM_PI * hu * hu * hu / 21.0));
        const float scale = std::ldexp(1.0f, e);
        auto *col = reinterpret_cast<unsigned long long *>(w->fl_list);
        constexpr bool kF = std::is_same<T, float>::value;
        // waves per chunk (the grid's waves loop over the items if fewer)
        const int parts = tuning().por_parts;
        if (chunked)
            hipLaunchKernelGGL((k_por_stream<T, true, kF, true>), dim3(blocks_for(n, bt)),
                               dim3(bt), 0, st, prm<T>(p), view<T>(P), view<T>(D), view<T>(gs),
                               view<T>(gd), (const T *)w->fl_V, w->fl_flag, cs, (T)hu, thr2, xt,
                               rm, col, scale, cl, parts, tuning().por_queue);
        else
            hipLaunchKernelGGL((k_por_stream<T, true, kF, false>), dim3(blocks_for(n, bt)),
                               dim3(bt), 0, st, prm<T>(p), view<T>(P), view<T>(D), view<T>(gs),
                               view<T>(gd), (const T *)w->fl_V, w->fl_flag, cs, (T)hu, thr2, xt,
                               rm, col, scale, cl, 1, 0);
        hipLaunchKernelGGL(k_por_finish<T>, dim3(blocks_for(n, bt)), dim3(bt), 0, st, prm<T>(p),
                           view<T>(P), view<T>(D), view<T>(gs), view<T>(gd), w->fl_flag,
                           (T *)w->fl_sum, col, std::ldexp(1.0, -e), cs, (T)hu, rm);
    } else {
        // FAST without the symmetric sums, or EXACT; with or without chunks
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(blocks_for(n, bt)), dim3(bt), 0, st, prm<T>(p),
                               view<T>(P), view<T>(D), view<T>(gs), view<T>(gd),
                               (const T *)w->fl_V, w->fl_flag, cs, (T)hu, thr2, xt, rm,
                               (unsigned long long *)nullptr, 0.0f, cl, 1, tuning().por_queue);
        };
        if (p->fast_math) {
            if (chunked) launch(k_por_stream<T, true, false, true>);
            else launch(k_por_stream<T, true, false, false>);
        } else {
            if (chunked) launch(k_por_stream<T, false, false, true>);
            else launch(k_por_stream<T, false, false, false>);
        }
    }
    return check_launch("porosity");
}
template int launch_porosity_tiled<float>(const sphdem_params_t *, const sphdem_sph_t *,
                                          const sphdem_dem_t *, const sphdem_grid_t *,
                                          const sphdem_grid_t *, const sphdem_work_t *, double,
                                          double, int *, hipStream_t, hipEvent_t);
template int launch_porosity_tiled<double>(const sphdem_params_t *, const sphdem_sph_t *,
                                           const sphdem_dem_t *, const sphdem_grid_t *,
                                           const sphdem_grid_t *, const sphdem_work_t *, double,
                                           double, int *, hipStream_t, hipEvent_t);

}  // namespace sphdem
