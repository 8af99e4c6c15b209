// Device helpers shared by the gfx950 kernels.  Functions marked
// MCV_HD compile for both host and device so that the host can check them
// against the real libstdc++ / libm they emulate.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MCV_HD __host__ __device__ __forceinline__
#else
#define MCV_HD inline
#endif

namespace mcv {

#if defined(__HIPCC__)
// Orders this wave's LDS accesses across its lanes (wave-private LDS areas).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b on XCD b % 8),
// each XCD with its own L2.  For a grid padded to a multiple of 8 this maps
// the physical block to a logical one so that every XCD works on one
// contiguous eighth of the logical blocks (neighbouring strips, cells or
// keypoints of the same image then share that XCD's L2); -1 for a padding
// block past the logical count nb.
__device__ __forceinline__ int xcd_block(int nb) {
    const int per = (int)gridDim.x >> 3;
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    return b < nb ? b : -1;
}

inline int xcd_grid(int nb) { return (nb + 7) / 8 * 8; }

// number of set bits of the wave mask b in the lanes below this one
__device__ __forceinline__ uint32_t lanes_below(unsigned long long b) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
}

// Ascending bitonic sort of n2 (a power of two) keys by the whole workgroup.
template <typename T, int kBatch = 8>
__device__ void bitonic_sort(T* a, int n2) {
    // each thread takes the compare-exchange pairs p = threadIdx.x + m *
    // blockDim.x of a stage (pair p of distance j starts at element
    // (p / j) * 2j + p % j); kBatch pairs are read before any is written, so
    // a stage costs one LDS round trip per kBatch pairs of a thread
    const int npairs = n2 >> 1;
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int p0 = 0; p0 < npairs; p0 += kBatch * (int)blockDim.x) {
                T x[kBatch], y[kBatch];
                int ia[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    const int p = p0 + u * (int)blockDim.x + (int)threadIdx.x;
                    ia[u] = ((p & ~(j - 1)) << 1) | (p & (j - 1));
                    if (p < npairs) {
                        x[u] = a[ia[u]];
                        y[u] = a[ia[u] | j];
                    }
                }
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    const int p = p0 + u * (int)blockDim.x + (int)threadIdx.x;
                    const bool up = (ia[u] & k) == 0;
                    if (p < npairs && (x[u] > y[u]) == up) {
                        a[ia[u]] = y[u];
                        a[ia[u] | j] = x[u];
                    }
                }
            }
            __syncthreads();
        }
    }
}
#endif

// ---------------------------------------------------------------------------
// std::priority_queue<T, vector<T>, Less> emulation on a plain array, with the
// exact sift order of libstdc++'s push_heap / pop_heap (bits/stl_heap.h:
// __push_heap, __adjust_heap, __pop_heap).  The ordering key of an entry is
// heap_key(); the rest of the entry is payload that never takes part in a
// comparison, exactly like the shared_ptr payload of the reference's node
// queue (ORBextractor.cc:530-531).  u64 entries keep the key in the high
// 32 bits; uint2 entries in .x >> 3 (the low 3 bits of .x are payload).
// Each step reads whole entries once (both children together on the way
// down), so a step costs one dependent memory round trip.
MCV_HD uint32_t heap_key(uint64_t a) { return (uint32_t)(a >> 32); }
#if defined(__HIPCC__)
MCV_HD uint32_t heap_key(const uint2& a) { return a.x >> 3; }
#endif

// `H` is any heap store with get(i) / set(i, v) (see ArrayHeap).
template <typename H, typename T>
MCV_HD void heap_sift_up(H& h, int hole, int top, T value) {
    while (hole > top) {
        const int parent = (hole - 1) / 2;
        const T p = h.get(parent);
        if (!(heap_key(p) < heap_key(value))) break;
        h.set(hole, p);
        hole = parent;
    }
    h.set(hole, value);
}

template <typename H, typename T>
MCV_HD void heap_push(H& h, int& len, T value) {
    heap_sift_up(h, len, 0, value);
    len++;
}

template <typename H, typename T>
MCV_HD void heap_adjust(H& h, int hole, int len, T value) {
    const int top = hole;
    int second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        const T r = h.get(second), l = h.get(second - 1);
        const bool take_left = heap_key(r) < heap_key(l);
        if (take_left) second--;
        h.set(hole, take_left ? l : r);
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        h.set(hole, h.get(second - 1));
        hole = second - 1;
    }
    heap_sift_up(h, hole, top, value);
}

template <typename H>
MCV_HD void heap_pop(H& h, int& len) {
    // libstdc++ also moves the old top to position len - 1, outside the
    // remaining heap; nothing reads it here, so that store is left out
    if (len > 1) {
        const int last = len - 1;
        heap_adjust(h, 0, last, h.get(last));
    }
    len--;
}

// ---------------------------------------------------------------------------
// std::sort(first, last, comp) with libstdc++'s exact element moves
// (bits/stl_algo.h: __introsort_loop with threshold 16 and depth limit
// 2 floor(log2 n), median-of-three to first, unguarded partition, heap sort
// at depth 0, final insertion sort), so that elements comparing equal end in
// the same order as with the reference's unstable sort.  `less(a, b)` is the
// comparator; one thread runs it on a small array.
template <typename T, typename Less>
MCV_HD void stl_unguarded_linear_insert(T* last, Less less) {
    T val = *last;
    T* next = last - 1;
    while (less(val, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = val;
}

template <typename T, typename Less>
MCV_HD void stl_insertion_sort(T* first, T* last, Less less) {
    if (first == last) return;
    for (T* i = first + 1; i != last; ++i) {
        if (less(*i, *first)) {
            T val = *i;
            for (T* p = i; p != first; --p) *p = *(p - 1);
            *first = val;
        } else {
            stl_unguarded_linear_insert(i, less);
        }
    }
}

template <typename T>
MCV_HD void stl_swap(T* a, T* b) {
    T t = *a;
    *a = *b;
    *b = t;
}

template <typename T, typename Less>
MCV_HD void stl_move_median_to_first(T* result, T* a, T* b, T* c, Less less) {
    if (less(*a, *b)) {
        if (less(*b, *c)) stl_swap(result, b);
        else if (less(*a, *c)) stl_swap(result, c);
        else stl_swap(result, a);
    } else if (less(*a, *c)) {
        stl_swap(result, a);
    } else if (less(*b, *c)) {
        stl_swap(result, c);
    } else {
        stl_swap(result, b);
    }
}

template <typename T, typename Less>
MCV_HD T* stl_unguarded_partition(T* first, T* last, T* pivot, Less less) {
    while (true) {
        while (less(*first, *pivot)) ++first;
        --last;
        while (less(*pivot, *last)) --last;
        if (!(first < last)) return first;
        stl_swap(first, last);
        ++first;
    }
}

// __adjust_heap / __make_heap / __pop_heap on [first, last) with `less`
template <typename T, typename Less>
MCV_HD void stl_adjust_heap(T* first, int hole, int len, T value, Less less) {
    const int top = hole;
    int second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (less(first[second], first[second - 1])) second--;
        first[hole] = first[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        first[hole] = first[second - 1];
        hole = second - 1;
    }
    int parent = (hole - 1) / 2;
    while (hole > top && less(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

template <typename T, typename Less>
MCV_HD void stl_heap_sort_all(T* first, T* last, Less less) {
    // __partial_sort(first, last, last): __heap_select (make_heap only, the
    // middle is the end) then __sort_heap
    const int len = (int)(last - first);
    if (len >= 2)
        for (int parent = (len - 2) / 2;; parent--) {
            stl_adjust_heap(first, parent, len, first[parent], less);
            if (parent == 0) break;
        }
    for (T* l = last; l - first > 1;) {
        --l;
        T value = *l;
        *l = *first;
        stl_adjust_heap(first, 0, (int)(l - first), value, less);
    }
}

// stack: room for 3 ints per pending range, 3 * (2 floor(log2 n) + 2) ints
// (each range pushed has a smaller depth budget than the one it came from);
// the caller places it (LDS on the device: no per-lane scratch)
template <typename T, typename Less>
MCV_HD void stl_sort(T* first, T* last, Less less, int* stack) {
    if (first == last) return;
    int n = (int)(last - first), lg = 0;
    while ((n >>= 1) != 0) lg++;
    // __introsort_loop, with the recursion on the right part as an explicit
    // stack of (first, last, depth) as offsets from `first`
    int sp = 0;
    stack[0] = 0;
    stack[1] = (int)(last - first);
    stack[2] = 2 * lg;
    sp = 1;
    while (sp > 0) {
        --sp;
        T* f = first + stack[3 * sp];
        T* l = first + stack[3 * sp + 1];
        int depth = stack[3 * sp + 2];
        while (l - f > 16) {
            if (depth == 0) {
                stl_heap_sort_all(f, l, less);
                break;
            }
            --depth;
            T* mid = f + (l - f) / 2;
            stl_move_median_to_first(f, f + 1, mid, l - 1, less);
            T* cut = stl_unguarded_partition(f + 1, l, f, less);
            // the reference recurses into [cut, l) before continuing with [f, cut)
            stack[3 * sp] = (int)(f - first);
            stack[3 * sp + 1] = (int)(cut - first);
            stack[3 * sp + 2] = depth;
            sp++;
            f = cut;
        }
        // (the pushed left parts are processed after the right parts, as the
        // reference's recursion order does)
    }
    // __final_insertion_sort
    if (last - first > 16) {
        stl_insertion_sort(first, first + 16, less);
        for (T* i = first + 16; i != last; ++i) stl_unguarded_linear_insert(i, less);
    } else {
        stl_insertion_sort(first, last, less);
    }
}

// A plain array as heap store.
template <typename T>
struct ArrayHeap {
    T* a;
    MCV_HD T get(int i) const { return a[i]; }
    MCV_HD void set(int i, T v) const { a[i] = v; }
};

// ---------------------------------------------------------------------------
// Base sin/cos of a float angle in [0, 8): Cody-Waite reduction by pi/2 in
// double, Taylor series to degree 19/20, one rounding to float.  Identical
// IEEE double arithmetic on host and device (no FMA contraction, no library
// calls), so the host can list every input where the result differs from
// the host libm's cosf/sinf; those inputs are patched from a table.
MCV_HD void sincos_base(float xf, float* c_out, float* s_out) {
    const double x = (double)xf;
    const double PIO2_HI = 1.5707963267948966;      // double(pi/2); its low 3 mantissa bits are 0
    const double PIO2_LO = 6.123233995736766e-17;   // pi/2 - PIO2_HI
    const double TWO_OVER_PI = 0.6366197723675814;
    int n = (int)(x * TWO_OVER_PI + 0.5);
    const double dn = (double)n;
    double r = x - dn * PIO2_HI;
    r = r - dn * PIO2_LO;
    const double r2 = r * r;
    // sin r = r - r^3 (1/3! - r^2/5! + ...), cos r = 1 - r^2 (1/2! - r^2/4! + ...)
    double s = 1.0 / 121645100408832000.0;    // 1/19!
    s = s * (-r2) + 1.0 / 355687428096000.0;  // 1/17!
    s = s * (-r2) + 1.0 / 1307674368000.0;    // 1/15!
    s = s * (-r2) + 1.0 / 6227020800.0;       // 1/13!
    s = s * (-r2) + 1.0 / 39916800.0;         // 1/11!
    s = s * (-r2) + 1.0 / 362880.0;           // 1/9!
    s = s * (-r2) + 1.0 / 5040.0;             // 1/7!
    s = s * (-r2) + 1.0 / 120.0;              // 1/5!
    s = s * (-r2) + 1.0 / 6.0;                // 1/3!
    s = r - (r * r2) * s;
    double c = 1.0 / 2432902008176640000.0;   // 1/20!
    c = c * (-r2) + 1.0 / 6402373705728000.0; // 1/18!
    c = c * (-r2) + 1.0 / 20922789888000.0;   // 1/16!
    c = c * (-r2) + 1.0 / 87178291200.0;      // 1/14!
    c = c * (-r2) + 1.0 / 479001600.0;        // 1/12!
    c = c * (-r2) + 1.0 / 3628800.0;          // 1/10!
    c = c * (-r2) + 1.0 / 40320.0;            // 1/8!
    c = c * (-r2) + 1.0 / 720.0;              // 1/6!
    c = c * (-r2) + 1.0 / 24.0;               // 1/4!
    c = c * (-r2) + 0.5;                      // 1/2!
    c = 1.0 - r2 * c;
    double cs, sn;
    switch (n & 3) {
        case 0: cs = c; sn = s; break;
        case 1: cs = -s; sn = c; break;
        case 2: cs = -c; sn = -s; break;
        default: cs = s; sn = -c; break;
    }
    *c_out = (float)cs;
    *s_out = (float)sn;
}

// ---------------------------------------------------------------------------
// cv::fastAtan2 in float (degrees); the constants are the float products
// 0.9997878412794807f * (float)(180/pi) etc. evaluated in float.
MCV_HD float fast_atan2_deg(float y, float x) {
    const float k = (float)(180.0 / 3.14159265358979323846);
    const float p1 = 0.9997878412794807f * k;
    const float p3 = -0.3258083974640975f * k;
    const float p5 = 0.1555786518463281f * k;
    const float p7 = -0.04432655554792128f * k;
    const float eps = (float)2.220446049250313080847e-16;  // (float)DBL_EPSILON
    float ax = x < 0 ? -x : x, ay = y < 0 ? -y : y;
    float a, c, c2;
    if (ax >= ay) {
        c = ay / (ax + eps);
        c2 = c * c;
        a = (((p7 * c2 + p5) * c2 + p3) * c2 + p1) * c;
    } else {
        c = ax / (ay + eps);
        c2 = c * c;
        a = 90.f - (((p7 * c2 + p5) * c2 + p3) * c2 + p1) * c;
    }
    if (x < 0) a = 180.f - a;
    if (y < 0) a = 360.f - a;
    return a;
}

}  // namespace mcv
