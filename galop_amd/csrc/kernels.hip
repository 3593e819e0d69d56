// kernels.hip -- gfx950 kernels of the Galop fitness evaluator.
//
//   plain_kernel       one program per wave64 over a CSR population; the
//                      replacement of evaluate_kernel_packed
//                      (src/EvaluatorCuda.cu:241-265) and, through the dense
//                      packing of the C ABI, of evaluate_kernel (:163-190).
//   run_cases_kernel   same interpreter, writes per-case outputs.
//   batched_kernel     persistent waves pulling programs from a work queue
//                      ordered longest-bucket-first; datasets that fit in one
//                      tile stay in VGPRs across programs.
//   effective_*        device intron removal (src/Introns.h:28-57).
//   scan_*             int64 exclusive scan for CSR offsets.
#include "galop_internal.h"

namespace galop {

// Where program p lives: CSR (words[offsets[p], offsets[p+1])), or, with
// dense_len, the engine's dense layout (words[p*stride, p*stride + len[p]));
// a dense length above the stride reads only the program's own slot.
__device__ __forceinline__ void program_bounds(const int64_t* __restrict__ offsets,
                                               const uint8_t* __restrict__ dense_len,
                                               int32_t stride, int64_t p, int64_t& lo,
                                               int64_t& hi) {
    if (dense_len) {
        lo = p * (int64_t)stride;
        hi = lo + min((int32_t)dense_len[p], stride);
    } else {
        lo = offsets[p];
        hi = offsets[p + 1];
    }
}

// The interpreter kernels ask for at least 4 waves per SIMD (at most 128
// VGPRs).  At 8 cases per lane the unrolled FP64 evaluations of a unary word
// otherwise take 130-140 VGPRs and 3 waves; 4 waves with a few spilled
// dwords measured 15% faster on the plain workload.  With 8 registers and at
// most 4 cases per lane they ask for 5 (at most 102 VGPRs), which the 4-case
// kernels need a few more VGPRs than to reach on their own; with 16
// registers 5 waves would spill hundreds of bytes.
#define GALOP_INTERP_BOUNDS \
    __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NR == 8 && K <= 4 ? 5 : 4)))

// ---------------------------------------------------------------------------
// plain kernel
// ---------------------------------------------------------------------------
// EFF: programs of at most kSkipMaxWords words run their effective words only
// (bitwise the same fitness).
template <int NR, int K, int EXT, int EFF>
__global__ void GALOP_INTERP_BOUNDS
plain_kernel(const uint32_t* __restrict__ words, const int64_t* __restrict__ offsets,
             const uint8_t* __restrict__ dense_len, int32_t stride, int64_t first, int64_t n,
             DevData ds, float* __restrict__ fitness, double* __restrict__ sse_out) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const int wave_in_block = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t p = first + (int64_t)blockIdx.x * (int64_t)(blockDim.x >> 6) + wave_in_block;
    if (p >= n) return;  // wave-uniform
    const ConstRegs cr = load_const_regs(ds.consts, lane);
    int64_t lo, hi;
    program_bounds(offsets, dense_len, stride, p, lo, hi);
    const double sse = canonical_sse(
        eval_program_wave<NR, K, EXT, EFF>(words + lo, hi - lo, ds, lane, cr, nullptr));
    if (lane == 0) {
        fitness[p] = fitness_of_sse(sse, ds);
        if (sse_out) sse_out[p] = sse;
    }
}

// Case-split evaluation, for populations too small to fill the GPU with one
// wave per program: wave w takes program w / splits and, with s = w % splits,
// the case blocks [s*bpw, min((s+1)*bpw, n_blocks)), writing one FP64 partial
// per block.  split_finish_kernel then sums each program's partials in block
// order, the same FP64 additions eval_program_wave makes, so the SSE and
// fitness are the bits the plain kernel gives.
template <int NR, int K, int EXT, int EFF>
__global__ void GALOP_INTERP_BOUNDS
split_kernel(const uint32_t* __restrict__ words, const int64_t* __restrict__ offsets,
             const uint8_t* __restrict__ dense_len, int32_t stride, int64_t n, int splits,
             int64_t bpw, DevData ds, double* __restrict__ partials) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const int wave_in_block = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t w = (int64_t)blockIdx.x * (int64_t)(blockDim.x >> 6) + wave_in_block;
    const int64_t p = w / splits;
    if (p >= n) return;  // wave-uniform
    const int64_t n_blocks = (ds.n_cases + kCaseBlock - 1) / kCaseBlock;
    const int64_t b0 = (w % splits) * bpw;
    const int64_t b1 = (b0 + bpw) < n_blocks ? (b0 + bpw) : n_blocks;
    const ConstRegs cr = load_const_regs(ds.consts, lane);
    const DestDecode dd = dest_decode(NR, ds.input_protect);
    int64_t lo, hi;
    program_bounds(offsets, dense_len, stride, p, lo, hi);
    if (EFF) {
        EffProgram e;
        if (hi - lo <= kSkipMaxWords)
            mark_effective<NR, EXT>(words + lo, hi - lo, lane, ds.out_reg, ds.input_protect, e);
        for (int64_t b = b0; b < b1; ++b) {
            const double part = block_sse_eff<NR, K, EXT>(words + lo, hi - lo, e, ds, lane, cr,
                                                          dd, b * kCaseBlock, nullptr);
            if (lane == 0) partials[p * n_blocks + b] = part;
        }
        return;
    }
    for (int64_t b = b0; b < b1; ++b) {
        const double part = block_sse<NR, K, EXT>(words + lo, hi - lo, ds, lane, cr, dd,
                                                  b * kCaseBlock, nullptr);
        if (lane == 0) partials[p * n_blocks + b] = part;
    }
}

__global__ void split_finish_kernel(const double* __restrict__ partials, int64_t n,
                                    int64_t n_blocks, DevData ds,
                                    float* __restrict__ fitness, double* __restrict__ sse_out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double sse = 0.0;
    for (int64_t b = 0; b < n_blocks; ++b) sse += partials[p * n_blocks + b];
    sse = canonical_sse(sse);
    fitness[p] = fitness_of_sse(sse, ds);
    if (sse_out) sse_out[p] = sse;
}

template <int NR, int EXT>
__global__ void __launch_bounds__(256)
run_cases_kernel(const uint32_t* __restrict__ words, const int64_t* __restrict__ offsets,
                 int64_t first, int64_t n, DevData ds, float* __restrict__ out) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const int wave_in_block = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t p = first + (int64_t)blockIdx.x * (int64_t)(blockDim.x >> 6) + wave_in_block;
    if (p >= n) return;
    const ConstRegs cr = load_const_regs(ds.consts, lane);
    const int64_t lo = offsets[p];
    const int64_t hi = offsets[p + 1];
    (void)eval_program_wave<NR, 1, EXT>(words + lo, hi - lo, ds, lane, cr, out + p * ds.n_cases);
}

__global__ void const_table_kernel(float* consts) {
    const int t = (int)threadIdx.x;  // 64 threads: unary op u on constant idx
    const int u = t / 16, idx = t % 16;
    const float b = consts[idx];
    const galop_math::CoefPtr C = galop_math::coef_table();
    float v;
    switch (u) {
    case 0: v = op_sin(b, C); break;
    case 1: v = op_cos(b, C); break;
    case 2: v = op_exp(b, C); break;
    default: v = op_plog(b, C); break;
    }
    consts[kUnaryTableBase + t] = v;
}

PlainTuning resolve_plain_tuning(const DevData& ds, int num_registers, PlainTuning t) {
    if (t.cases_per_lane <= 0) {
        // enough cases per lane to amortise the scalar decode and dispatch,
        // without padding small datasets: K = ceil(N/64) capped at 8 (8
        // registers) / 4 (16 registers)
        const int64_t per_lane = (ds.n_cases + kWave - 1) / kWave;
        int kmax = num_registers == 8 ? 8 : 4;
        int k = 1;
        while (k < kmax && k < per_lane) k *= 2;
        t.cases_per_lane = k;
    }
    // 16 registers x 8 cases would exceed the 32-float register groups x 4
    if (num_registers == 16 && t.cases_per_lane > 4) t.cases_per_lane = 4;
    if (t.waves_per_block <= 0 || t.waves_per_block > 4) t.waves_per_block = 4;
    return t;
}

// Case splitting pays when the population leaves the GPU short of waves and
// each program has several case blocks.  Returns the waves per program (1 =
// no split) and sets *bpw to the blocks each split wave takes.
static int plain_splits(const DevData& ds, int64_t n, int n_cu, int64_t* bpw) {
    const int64_t n_blocks = (ds.n_cases + kCaseBlock - 1) / kCaseBlock;
    const int64_t target = (int64_t)n_cu * 16;  // 4 waves on every SIMD
    *bpw = n_blocks;
    if (n_blocks <= 1 || n <= 0 || n >= target) return 1;
    int64_t splits = (target + n - 1) / n;
    if (splits > n_blocks) splits = n_blocks;
    *bpw = (n_blocks + splits - 1) / splits;
    return (int)((n_blocks + *bpw - 1) / *bpw);
}

int64_t plain_split_partials(const DevData& ds, int64_t n, int n_cu) {
    int64_t bpw;
    if (plain_splits(ds, n, n_cu, &bpw) <= 1) return 0;
    return n * ((ds.n_cases + kCaseBlock - 1) / kCaseBlock);
}

template <int NR, int K, int EXT, int EFF>
static hipError_t launch_plain_t(const DevData& ds, int wpb, const uint32_t* words,
                                 const int64_t* offsets, const uint8_t* dense_len,
                                 int32_t stride, int64_t n, float* fitness, double* sse,
                                 double* partials, int n_cu, hipStream_t stream) {
    int64_t bpw = 0;
    const int splits = partials ? plain_splits(ds, n, n_cu, &bpw) : 1;
    if (splits > 1) {
        const int64_t waves = n * splits;
        hipLaunchKernelGGL((split_kernel<NR, K, EXT, EFF>), dim3((unsigned)((waves + wpb - 1) / wpb)),
                           dim3(wpb * kWave), 0, stream, words, offsets, dense_len, stride, n,
                           splits, bpw, ds, partials);
        hipLaunchKernelGGL(split_finish_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           stream, partials, n, (ds.n_cases + kCaseBlock - 1) / kCaseBlock,
                           ds, fitness, sse);
        return hipGetLastError();
    }
    for (int64_t first = 0; first < n; first += kLaunchWindow) {
        const int64_t cnt = (n - first) < kLaunchWindow ? (n - first) : kLaunchWindow;
        const unsigned grid = (unsigned)((cnt + wpb - 1) / wpb);
        hipLaunchKernelGGL((plain_kernel<NR, K, EXT, EFF>), dim3(grid), dim3(wpb * kWave), 0, stream,
                           words, offsets, dense_len, stride, first, first + cnt, ds, fitness,
                           sse);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_plain(const DevData& ds, int num_registers, PlainTuning t,
                        const uint32_t* words, const int64_t* offsets, int64_t n,
                        float* fitness, double* sse, double* partials, int n_cu,
                        hipStream_t stream, const uint8_t* dense_len, int32_t stride,
                        bool effective_only) {
    if (n <= 0) return hipSuccess;
    t = resolve_plain_tuning(ds, num_registers, t);
    // the effective-word variant keeps four chunk registers and the masks
    // live: at the largest tile of each register count it would spill the
    // register file, so it runs the next tile down
    if (effective_only) t.cases_per_lane = std::min(t.cases_per_lane, num_registers == 8 ? 4 : 2);
    const int wpb = t.waves_per_block;
#define GALOP_ARGS ds, wpb, words, offsets, dense_len, stride, n, fitness, sse, partials, n_cu, stream
#define GALOP_CASE(NRV, KV)                                                            \
    if (num_registers == NRV && t.cases_per_lane == KV && !effective_only)             \
        return ds.op_set ? launch_plain_t<NRV, KV, 1, 0>(GALOP_ARGS)                   \
                         : launch_plain_t<NRV, KV, 0, 0>(GALOP_ARGS);
#define GALOP_CASE_EFF(NRV, KV)                                                        \
    if (num_registers == NRV && t.cases_per_lane == KV && effective_only)              \
        return ds.op_set ? launch_plain_t<NRV, KV, 1, 1>(GALOP_ARGS)                   \
                         : launch_plain_t<NRV, KV, 0, 1>(GALOP_ARGS);
    GALOP_CASE(8, 1) GALOP_CASE(8, 2) GALOP_CASE(8, 4) GALOP_CASE(8, 8)
    GALOP_CASE(16, 1) GALOP_CASE(16, 2) GALOP_CASE(16, 4)
    GALOP_CASE_EFF(8, 1) GALOP_CASE_EFF(8, 2) GALOP_CASE_EFF(8, 4)
    GALOP_CASE_EFF(16, 1) GALOP_CASE_EFF(16, 2)
#undef GALOP_CASE
#undef GALOP_CASE_EFF
#undef GALOP_ARGS
    return hipErrorInvalidValue;
}

hipError_t launch_run_cases(const DevData& ds, int num_registers, const uint32_t* words,
                            const int64_t* offsets, int64_t n, float* out,
                            hipStream_t stream) {
    const int wpb = 4;
    for (int64_t first = 0; first < n; first += kLaunchWindow) {
        const int64_t cnt = (n - first) < kLaunchWindow ? (n - first) : kLaunchWindow;
        const unsigned grid = (unsigned)((cnt + wpb - 1) / wpb);
        const dim3 g(grid), b(wpb * kWave);
        if (num_registers == 8 && ds.op_set)
            hipLaunchKernelGGL((run_cases_kernel<8, 1>), g, b, 0, stream, words, offsets, first, first + cnt, ds, out);
        else if (num_registers == 8)
            hipLaunchKernelGGL((run_cases_kernel<8, 0>), g, b, 0, stream, words, offsets, first, first + cnt, ds, out);
        else if (ds.op_set)
            hipLaunchKernelGGL((run_cases_kernel<16, 1>), g, b, 0, stream, words, offsets, first, first + cnt, ds, out);
        else
            hipLaunchKernelGGL((run_cases_kernel<16, 0>), g, b, 0, stream, words, offsets, first, first + cnt, ds, out);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_const_table(float* consts, hipStream_t stream) {
    hipLaunchKernelGGL(const_table_kernel, dim3(1), dim3(64), 0, stream, consts);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// device intron removal (src/Introns.h:28-57): one thread per program walks
// its program backwards with a live-register bitmask.  With an order array
// (the bucket order of the batched path) thread i takes program order[i], so
// the lanes of a wave walk programs of similar length, and the counts and the
// compacted programs are laid out by order position i.
// ---------------------------------------------------------------------------
struct DecodeCfg {
    uint32_t mask;        // NR - 1
    uint32_t dmask;       // the destination field's mask: NR/2 - 1 under INPUT_PROTECT
    int dbase;            // added to the destination field: NR/2 under INPUT_PROTECT
    uint32_t unary_mask;  // bit op set for the unary opcodes of the op table
    int out_reg;
};

__device__ __forceinline__ bool effective_step(uint32_t w, const DecodeCfg& c, uint32_t& live) {
    const int d = (int)((w >> 8) & c.dmask) + c.dbase;
    if (!((live >> d) & 1u)) return false;
    live &= ~(1u << d);
    const bool unary = (c.unary_mask >> (w & 7u)) & 1u;
    if (!unary) live |= 1u << ((w >> 16) & c.mask);
    if (!(w >> 31)) live |= 1u << ((w >> 24) & c.mask);
    return true;
}

// Calls f(word) for the words of [lo, hi), last to first, reading the
// 128-byte aligned segments that hold them, one whole segment (eight 16-byte
// loads) per step.  The lanes of a wave walk 64 programs at scattered
// addresses; taking a cache line's worth per step fetches each line once,
// where a 16-byte piece per step held every line across eight steps, more
// lines than the caches keep, so lines were fetched again.  A segment never
// crosses a page, so its other words are readable whenever one in range is;
// they are not used.
template <typename F>
__device__ __forceinline__ void walk_backward(const uint32_t* __restrict__ words, int64_t lo,
                                              int64_t hi, F f) {
    if (hi <= lo) return;
    constexpr int kSeg = 32;  // words per segment
    const uintptr_t last = (uintptr_t)(words + (hi - 1));
    uintptr_t a = last & ~(uintptr_t)(4 * kSeg - 1);
    int64_t seg = (hi - 1) - (int64_t)((last - a) >> 2);  // word index of the segment's first word
    int top = (int)(hi - seg);                            // words below hi in this segment, 1..32
    for (;;) {
        const int bot = seg >= lo ? 0 : (int)(lo - seg);  // words below lo in this segment
        // bit t set for the segment's words t in [bot, top): one 32-bit test
        // per word (a comparison of t with the bounds was done on 64-bit
        // word indices)
        const uint32_t take = (top >= kSeg ? ~0u : (1u << top) - 1u) & ~((1u << bot) - 1u);
        // global loads (a flat load would also wait on the LDS counter)
        typedef uint32_t U32x4 __attribute__((ext_vector_type(4)));
        typedef const __attribute__((address_space(1))) U32x4* GlobalBlocks;
        U32x4 v[kSeg / 4];
#pragma unroll
        for (int q = 0; q < kSeg / 4; ++q) v[q] = reinterpret_cast<GlobalBlocks>(a)[q];
#pragma unroll
        for (int t = kSeg - 1; t >= 0; --t) {
            const uint32_t w = v[t / 4][t & 3];
            if (take & (1u << t)) f(w);
        }
        if (seg <= lo) break;
        seg -= kSeg;
        a -= 4 * kSeg;
        top = kSeg;
    }
}

__global__ void effective_count_kernel(DecodeCfg cfg, const uint32_t* __restrict__ words,
                                       const int64_t* __restrict__ offsets,
                                       const int32_t* __restrict__ order, int64_t n,
                                       int64_t* __restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = order ? (int64_t)order[i] : i;
    uint32_t live = 1u << cfg.out_reg;
    int64_t cnt = 0;
    walk_backward(words, offsets[p], offsets[p + 1],
                  [&](uint32_t w) { cnt += effective_step(w, cfg, live) ? 1 : 0; });
    counts[i] = cnt;
}

// Writes a program's effective words from the back of its output range
// [end - count, end): the walk finds them last to first, and the staging block
// holds the last four written, s_j at position end - cnt + j.  When a word
// lands at a position divisible by four, the four positions from it up are
// complete and go out with one 16-byte store (word by word where the block
// reaches `lim`, the end of what may be written, or the buffer is not 16-byte
// aligned); after the walk, finish() writes the partial block at the front.
// Only positions of this program's own effective words below `lim` are ever
// written.
struct BackWriter {
    uint32_t* __restrict__ pend;  // out + end: position end - c is pend - c
    int over;                     // positions end - c with c <= over lie past lim
    uint32_t ealign;              // end & 3: position end - c is a multiple of 4 iff (c & 3) == ealign
    bool vec_ok;
    int cnt = 0;  // words written so far (a program has fewer than 2^31 words)
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;

    __device__ BackWriter(uint32_t* o, int64_t e, int64_t l, bool v)
        : pend(o + e), over((int)(e - l < (int64_t(1) << 30) ? e - l : (int64_t(1) << 30))),
          ealign((uint32_t)e & 3u), vec_ok(v) {}
    // positions end - c .. end - c + n - 1 from s0 .. s(n-1), none at or past
    // lim (position end - c + j is below lim iff c - j > over)
    __device__ __forceinline__ void store(int c, int n) {
        uint32_t* p = pend - c;
        if (vec_ok && n == 4 && c - 4 >= over) {
            *reinterpret_cast<uint4*>(p) = make_uint4(s0, s1, s2, s3);
            return;
        }
        if (n > 0 && c > over) p[0] = s0;
        if (n > 1 && c - 1 > over) p[1] = s1;
        if (n > 2 && c - 2 > over) p[2] = s2;
        if (n > 3 && c - 3 > over) p[3] = s3;
    }
    __device__ __forceinline__ void push(uint32_t w) {
        s3 = s2;
        s2 = s1;
        s1 = s0;
        s0 = w;
        ++cnt;
        // s_j for j >= cnt are not this program's: their positions are at or
        // past end, and store() writes none of them
        if (((uint32_t)cnt & 3u) == ealign) store(cnt, 4);
    }
    __device__ __forceinline__ void finish() {
        const int r = (int)((ealign - (uint32_t)cnt) & 3u);  // (end - cnt) & 3
        if (cnt > 0 && r != 0) store(cnt, min(4 - r, cnt));
    }
};

// The compacted program i goes to [eff_offsets[i], eff_offsets[i + 1]) of
// eff_words, written from the back (BackWriter).
__global__ void effective_scatter_kernel(DecodeCfg cfg, const uint32_t* __restrict__ words,
                                         const int64_t* __restrict__ offsets,
                                         const int32_t* __restrict__ order, int64_t n,
                                         const int64_t* __restrict__ eff_offsets,
                                         uint32_t* __restrict__ eff_words, int64_t cap) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = order ? (int64_t)order[i] : i;
    const int64_t end = eff_offsets[i + 1];
    BackWriter out(eff_words, end, end < cap ? end : cap,  // never write past the buffer
                   ((uintptr_t)eff_words & 15) == 0);
    uint32_t live = 1u << cfg.out_reg;
    walk_backward(words, offsets[p], offsets[p + 1], [&](uint32_t w) {
        if (effective_step(w, cfg, live)) out.push(w);
    });
    out.finish();
}

// Intron removal for the batched evaluation in one pass, without a prefix
// sum: thread i walks program p = order[i] backwards and writes its effective
// words backwards from the end of p's own raw slot, in a buffer laid out like
// the raw CSR (positions relative to offsets[0]).  A program has no more
// effective words than raw ones, so they stay inside its slot: its effective
// code is slot[end_p - counts[i], end_p).  Writes go through BackWriter, as
// in effective_scatter_kernel: only the positions of p's effective words below
// min(end_p, cap) are written.
__global__ void effective_in_slot_kernel(DecodeCfg cfg, const uint32_t* __restrict__ words,
                                         const int64_t* __restrict__ offsets,
                                         const int32_t* __restrict__ order, int64_t n,
                                         uint32_t* __restrict__ slot, int64_t cap,
                                         int64_t* __restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = order[i];
    const int64_t end = offsets[p + 1] - offsets[0];
    BackWriter out(slot, end, end < cap ? end : cap, ((uintptr_t)slot & 15) == 0);
    uint32_t live = 1u << cfg.out_reg;
    walk_backward(words, offsets[p], offsets[p + 1], [&](uint32_t w) {
        if (effective_step(w, cfg, live)) out.push(w);
    });
    out.finish();
    counts[i] = out.cnt;
}

// Offsets past the effective-word buffer are cut to its end, so that a
// population larger than the buffer can never make a kernel read past it.
__global__ void clamp_offsets_kernel(int64_t* offsets, int64_t n1, int64_t cap) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n1 && offsets[i] > cap) offsets[i] = cap;
}

static DecodeCfg make_cfg(int nr, int op_set, int iprot) {
    DecodeCfg c;
    c.mask = (uint32_t)nr - 1u;
    c.dmask = iprot ? (uint32_t)(nr / 2 - 1) : c.mask;
    c.dbase = iprot ? nr / 2 : 0;
    // SIN, COS (src/ISA.h:209-212,225-226: the first operand is ignored); exp and log
    // in the EXT8 table
    c.unary_mask = 0x30u | (op_set ? 0xc0u : 0u);
    c.out_reg = iprot ? nr / 2 : 0;
    return c;
}

hipError_t launch_effective_in_slot(int num_registers, int op_set, int input_protect,
                                    const uint32_t* words, const int64_t* offsets, int64_t n,
                                    const int32_t* order, uint32_t* slot, int64_t cap,
                                    int64_t* counts, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(effective_in_slot_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       stream, make_cfg(num_registers, op_set, input_protect), words, offsets,
                       order, n, slot, cap, counts);
    return hipGetLastError();
}

hipError_t launch_effective_counts(int num_registers, int op_set, int input_protect,
                                   const uint32_t* words, const int64_t* offsets, int64_t n,
                                   int64_t* counts, hipStream_t stream, const int32_t* order) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(effective_count_kernel, dim3(grid), dim3(256), 0, stream,
                       make_cfg(num_registers, op_set, input_protect), words, offsets, order, n,
                       counts);
    return hipGetLastError();
}

hipError_t launch_effective_scatter(int num_registers, int op_set, int input_protect,
                                    const uint32_t* words, const int64_t* offsets, int64_t n,
                                    int64_t* eff_offsets, uint32_t* eff_words, int64_t cap,
                                    hipStream_t stream, const int32_t* order) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(effective_scatter_kernel, dim3(grid), dim3(256), 0, stream,
                       make_cfg(num_registers, op_set, input_protect), words, offsets, order, n,
                       eff_offsets, eff_words, cap);
    hipLaunchKernelGGL(clamp_offsets_kernel, dim3((unsigned)((n + 256) / 256)), dim3(256), 0, stream,
                       eff_offsets, n + 1, cap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// int64 exclusive scan: block sums -> one-block scan of the sums -> rescan.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int64_t kScanTile = (int64_t)kScanThreads * kScanItems;

__device__ int64_t block_exclusive_scan(int64_t v, int64_t* sh, int64_t& total) {
    // Hillis-Steele over kScanThreads entries in LDS
    const int t = (int)threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int64_t x = (t >= off) ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    total = sh[kScanThreads - 1];
    const int64_t incl = sh[t];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(kScanThreads)
scan_block_sums(const int64_t* __restrict__ in, int64_t n, int64_t* __restrict__ sums) {
    __shared__ int64_t sh[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t s = 0;
    for (int k = 0; k < kScanItems; ++k) s += (base + k < n) ? in[base + k] : 0;
    int64_t total;
    (void)block_exclusive_scan(s, sh, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads)
scan_sums_single(int64_t* __restrict__ sums, int64_t nb) {
    __shared__ int64_t sh[kScanThreads];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += kScanThreads) {
        const int64_t i = base + threadIdx.x;
        const int64_t v = (i < nb) ? sums[i] : 0;
        int64_t total;
        const int64_t ex = block_exclusive_scan(v, sh, total);
        if (i < nb) sums[i] = carry + ex;
        carry += total;
    }
}

__global__ void __launch_bounds__(kScanThreads)
scan_apply(const int64_t* __restrict__ in, int64_t n, const int64_t* __restrict__ sums,
           int64_t* __restrict__ out) {
    __shared__ int64_t sh[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int64_t total;
    int64_t run = sums[blockIdx.x] + block_exclusive_scan(s, sh, total);
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (base <= n - 1 && n - 1 < base + kScanItems) out[n] = run;  // total after the last item
}

__global__ void scan_zero_kernel(int64_t* out) { out[0] = 0; }

int64_t scan_temp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

hipError_t launch_exclusive_scan(const int64_t* counts, int64_t n, int64_t* out, int64_t* tmp,
                                 hipStream_t stream) {
    if (n <= 0) {  // out[0] = 0 (a kernel, not a memset, so the call can be captured)
        hipLaunchKernelGGL(scan_zero_kernel, dim3(1), dim3(1), 0, stream, out);
        return hipGetLastError();
    }
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(scan_block_sums, dim3((unsigned)nb), dim3(kScanThreads), 0, stream,
                       counts, n, tmp);
    hipLaunchKernelGGL(scan_sums_single, dim3(1), dim3(kScanThreads), 0, stream, tmp, nb);
    hipLaunchKernelGGL(scan_apply, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, counts, n,
                       tmp, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// batched persistent path
// ---------------------------------------------------------------------------
// Length bucket: 0..15 exact for lengths 0..15, then 4 buckets per octave
// up to bucket 63 (lengths >= ~2^15 share the last bucket).
__device__ __forceinline__ int length_bucket(int64_t len) {
    if (len < 16) return (int)len;
    const int lg = 63 - __clzll((unsigned long long)len);       // >= 4
    const int frac = (int)((len >> (lg - 2)) & 3);               // next 2 bits
    const int b = 16 + (lg - 4) * 4 + frac;
    return b > kBuckets - 1 ? kBuckets - 1 : b;
}

__global__ void batch_reset_kernel(int32_t* bucket_count, int32_t* bucket_cursor,
                                   int32_t* queue_head) {
    const int t = (int)threadIdx.x;
    if (t < kBuckets) {
        bucket_count[t] = 0;
        bucket_cursor[t] = 0;
    }
    if (t == 0) *queue_head = 0;
}

__global__ void batch_histogram_kernel(const int64_t* __restrict__ offsets, int64_t n,
                                       int32_t* __restrict__ bucket_count) {
    __shared__ int32_t local[kBuckets];
    if (threadIdx.x < kBuckets) local[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&local[length_bucket(offsets[i + 1] - offsets[i])], 1);
    __syncthreads();
    if (threadIdx.x < kBuckets && local[threadIdx.x])
        atomicAdd(&bucket_count[threadIdx.x], local[threadIdx.x]);
}

// Start of each bucket in the order array, longest bucket first.
__global__ void batch_bucket_start_kernel(const int32_t* __restrict__ bucket_count,
                                          int32_t* __restrict__ bucket_cursor) {
    if (threadIdx.x == 0) {
        int32_t run = 0;
        for (int b = kBuckets - 1; b >= 0; --b) {
            bucket_cursor[b] = run;
            run += bucket_count[b];
        }
    }
}

// Order array by bucket.  Each block ranks its tile of programs inside the
// tile with LDS atomics, then reserves one range per bucket with a single
// global atomic, so the global counters see at most kBuckets atomics per
// tile rather than one per program.
constexpr int kScatterThreads = 256;
constexpr int kScatterItems = 16;
constexpr int64_t kScatterTile = (int64_t)kScatterThreads * kScatterItems;

__global__ void __launch_bounds__(kScatterThreads)
batch_scatter_kernel(const int64_t* __restrict__ offsets, int64_t n,
                     int32_t* __restrict__ bucket_cursor, int32_t* __restrict__ order) {
    __shared__ int32_t cnt[kBuckets];
    __shared__ int32_t base[kBuckets];
    const int t = (int)threadIdx.x;
    if (t < kBuckets) cnt[t] = 0;
    __syncthreads();
    const int64_t first = (int64_t)blockIdx.x * kScatterTile;
    int bucket[kScatterItems], rank[kScatterItems];
#pragma unroll
    for (int it = 0; it < kScatterItems; ++it) {
        const int64_t i = first + (int64_t)it * kScatterThreads + t;
        bucket[it] = 0;
        rank[it] = 0;
        if (i < n) {
            bucket[it] = length_bucket(offsets[i + 1] - offsets[i]);
            rank[it] = atomicAdd(&cnt[bucket[it]], 1);
        }
    }
    __syncthreads();
    if (t < kBuckets && cnt[t]) base[t] = atomicAdd(&bucket_cursor[t], cnt[t]);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kScatterItems; ++it) {
        const int64_t i = first + (int64_t)it * kScatterThreads + t;
        if (i < n) order[base[bucket[it]] + rank[it]] = (int32_t)i;
    }
}

__device__ __forceinline__ int64_t readlane_i64(int64_t v, int j) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), j);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Persistent evaluation: each wave repeatedly takes `chunk` (<= 64) entries of
// the order array from the queue head, loads their program ids and CSR
// bounds one entry per lane, and evaluates them one after the other.  With
// by_position the CSR is indexed by order position (the compacted programs of
// the intron-removing generation), otherwise by program id.  Every wave
// leaves the loop once the head passes n, so the grid always drains.
template <int NR, int K, int EXT, int STAGE>
__global__ void __launch_bounds__(256)
batched_kernel(const uint32_t* __restrict__ words, const int64_t* __restrict__ offs,
               const int32_t* __restrict__ order, int32_t* __restrict__ queue_head,
               int64_t n, int chunk, const int64_t* __restrict__ eff_counts, int64_t cap,
               DevData ds, float* __restrict__ fitness,
               double* __restrict__ sse_out) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const ConstRegs cr = load_const_regs(ds.consts, lane);
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(queue_head, chunk);
        base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, kWave));
        if ((int64_t)base >= n) break;
        const int cnt = (int)(((int64_t)base + chunk) < n ? chunk : (n - base));
        int32_t p_l = 0;
        int64_t lo_l = 0, hi_l = 0;
        if (lane < cnt) {
            const int64_t q = (int64_t)base + lane;
            p_l = order[q];
            if (eff_counts) {
                // effective code at the end of the program's raw slot
                // (effective_in_slot_kernel), cut at the buffer's end
                const int64_t end = offs[p_l + 1] - offs[0];
                hi_l = end < cap ? end : cap;
                lo_l = end - eff_counts[q];
                lo_l = lo_l < hi_l ? lo_l : hi_l;
            } else {
                lo_l = offs[p_l];
                hi_l = offs[p_l + 1];
            }
        }
        for (int j = 0; j < cnt; ++j) {
            const int64_t p = __builtin_amdgcn_readlane(p_l, j);
            const int64_t lo = readlane_i64(lo_l, j);
            const int64_t hi = readlane_i64(hi_l, j);
            const double sse = canonical_sse(
                eval_program_wave<NR, K, EXT, 0, STAGE>(words + lo, hi - lo, ds, lane, cr, nullptr));
            if (lane == 0) {
                fitness[p] = fitness_of_sse(sse, ds);
                if (sse_out) sse_out[p] = sse;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// serving rounds: the programs of every request of a round in one work queue
// ---------------------------------------------------------------------------
__device__ __forceinline__ int work_bucket(int64_t work) {
    if (work < 16) return (int)(work < 0 ? 0 : work);
    const int lg = 63 - __clzll((unsigned long long)work);  // >= 4
    const int frac = (int)((work >> (lg - 2)) & 3);
    const int b = 16 + (lg - 4) * 4 + frac;
    return b > kServeBuckets - 1 ? kServeBuckets - 1 : b;
}

// counters: [0, B) program counts per bucket, [B, 2B) bucket starts, [2B, 3B)
// places taken so far in each bucket (B = kServeBuckets).
__global__ void serve_reset_kernel(int32_t* counters) {
    for (int i = (int)threadIdx.x; i < 3 * kServeBuckets; i += (int)blockDim.x) counters[i] = 0;
}

__global__ void __launch_bounds__(256)
serve_histogram_kernel(const ServeDesc* __restrict__ desc, int r, int32_t* __restrict__ counts) {
    __shared__ int32_t local[kServeBuckets];
    for (int i = (int)threadIdx.x; i < kServeBuckets; i += (int)blockDim.x) local[i] = 0;
    __syncthreads();
    const int64_t* __restrict__ off = desc[r].offsets;
    const int64_t n = desc[r].n, tiles = desc[r].tiles;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&local[work_bucket((off[p + 1] - off[p]) * tiles)], 1);
    __syncthreads();
    for (int i = (int)threadIdx.x; i < kServeBuckets; i += (int)blockDim.x)
        if (local[i]) atomicAdd(&counts[i], local[i]);
}

// Start of each bucket in the queue, heaviest bucket first.
__global__ void serve_bucket_start_kernel(const int32_t* __restrict__ counts,
                                          int32_t* __restrict__ cursor) {
    if (threadIdx.x == 0) {
        int32_t run = 0;
        for (int b = kServeBuckets - 1; b >= 0; --b) {
            cursor[b] = run;
            run += counts[b];
        }
    }
}

// Queue entries of request r: each block ranks its tile of programs in LDS and
// reserves one range per bucket with a single global atomic.
__global__ void __launch_bounds__(kScatterThreads)
serve_scatter_kernel(const ServeDesc* __restrict__ desc, int r, const int32_t* __restrict__ start,
                     int32_t* __restrict__ taken, int64_t* __restrict__ order) {
    __shared__ int32_t cnt[kServeBuckets];
    __shared__ int32_t base[kServeBuckets];
    const int t = (int)threadIdx.x;
    for (int i = t; i < kServeBuckets; i += kScatterThreads) cnt[i] = 0;
    __syncthreads();
    const int64_t* __restrict__ off = desc[r].offsets;
    const int64_t n = desc[r].n, tiles = desc[r].tiles;
    const int64_t first = (int64_t)blockIdx.x * kScatterTile;
    int bucket[kScatterItems], rank[kScatterItems];
#pragma unroll
    for (int it = 0; it < kScatterItems; ++it) {
        const int64_t p = first + (int64_t)it * kScatterThreads + t;
        bucket[it] = 0;
        rank[it] = 0;
        if (p < n) {
            bucket[it] = work_bucket((off[p + 1] - off[p]) * tiles);
            rank[it] = atomicAdd(&cnt[bucket[it]], 1);
        }
    }
    __syncthreads();
    for (int i = t; i < kServeBuckets; i += kScatterThreads)
        if (cnt[i]) base[i] = start[i] + atomicAdd(&taken[i], cnt[i]);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kScatterItems; ++it) {
        const int64_t p = first + (int64_t)it * kScatterThreads + t;
        if (p < n)
            order[(int64_t)base[bucket[it]] + rank[it]] = ((int64_t)r << kServeProgramBits) | p;
    }
}

// One wave per queue entry: the request's descriptor is read with the
// wave-uniform request index, then the program runs exactly as in
// plain_kernel (same bits).
template <int NR, int K, int EXT>
__global__ void GALOP_INTERP_BOUNDS
serve_kernel(const ServeDesc* __restrict__ desc, const int64_t* __restrict__ order,
             int64_t first, int64_t n) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const int wave_in_block = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t w = first + (int64_t)blockIdx.x * (int64_t)(blockDim.x >> 6) + wave_in_block;
    if (w >= n) return;  // wave-uniform
    const int64_t e = readlane_i64(order[w], 0);
    const int r = (int)(e >> kServeProgramBits);
    const int64_t p = e & ((int64_t(1) << kServeProgramBits) - 1);
    const ServeDesc* __restrict__ d = desc + r;
    const DevData ds = d->ds;
    const ConstRegs cr = load_const_regs(ds.consts, lane);
    // the bounds are wave-uniform: broadcast, so the program's base and
    // length live in SGPRs
    const int64_t lo = readlane_i64(d->offsets[p], 0), hi = readlane_i64(d->offsets[p + 1], 0);
    const double sse =
        canonical_sse(eval_program_wave<NR, K, EXT>(d->words + lo, hi - lo, ds, lane, cr, nullptr));
    if (lane == 0) {
        d->fitness[p] = fitness_of_sse(sse, ds);
        if (d->sse) d->sse[p] = sse;
    }
}

hipError_t launch_serve_order(const ServeDesc* desc, const ServeDesc* h_desc, int n_req,
                              int32_t* counters, int64_t* order, hipStream_t stream) {
    hipLaunchKernelGGL(serve_reset_kernel, dim3(1), dim3(256), 0, stream, counters);
    for (int r = 0; r < n_req; ++r) {
        const int64_t n = h_desc[r].n;
        if (n <= 0) continue;
        const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
        hipLaunchKernelGGL(serve_histogram_kernel, dim3(grid), dim3(256), 0, stream, desc, r,
                           counters);
    }
    hipLaunchKernelGGL(serve_bucket_start_kernel, dim3(1), dim3(64), 0, stream, counters,
                       counters + kServeBuckets);
    for (int r = 0; r < n_req; ++r) {
        const int64_t n = h_desc[r].n;
        if (n <= 0) continue;
        hipLaunchKernelGGL(serve_scatter_kernel, dim3((unsigned)((n + kScatterTile - 1) / kScatterTile)),
                           dim3(kScatterThreads), 0, stream, desc, r, counters + kServeBuckets,
                           counters + 2 * kServeBuckets, order);
    }
    return hipGetLastError();
}

template <int NR, int K, int EXT>
static hipError_t launch_serve_t(const ServeDesc* desc, const int64_t* order, int64_t n,
                                 hipStream_t stream) {
    constexpr int wpb = 4;
    for (int64_t first = 0; first < n; first += kLaunchWindow) {
        const int64_t cnt = (n - first) < kLaunchWindow ? (n - first) : kLaunchWindow;
        hipLaunchKernelGGL((serve_kernel<NR, K, EXT>), dim3((unsigned)((cnt + wpb - 1) / wpb)),
                           dim3(wpb * kWave), 0, stream, desc, order, first, first + cnt);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_serve_eval(const ServeDesc* desc, int num_registers, int cases_per_lane,
                             int op_set, const int64_t* order, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
#define GALOP_CASE(NRV, KV)                                                                 \
    if (num_registers == NRV && cases_per_lane == KV)                                       \
        return op_set ? launch_serve_t<NRV, KV, 1>(desc, order, n, stream)                  \
                      : launch_serve_t<NRV, KV, 0>(desc, order, n, stream);
    GALOP_CASE(8, 1) GALOP_CASE(8, 2) GALOP_CASE(8, 4) GALOP_CASE(8, 8)
    GALOP_CASE(16, 1) GALOP_CASE(16, 2) GALOP_CASE(16, 4)
#undef GALOP_CASE
    return hipErrorInvalidValue;
}

BatchTuning resolve_batch_tuning(const DevData& ds, int num_registers, BatchTuning t, int n_cu) {
    (void)n_cu;
    if (t.cases_per_lane <= 0) {
        PlainTuning pt;
        pt = resolve_plain_tuning(ds, num_registers, pt);
        t.cases_per_lane = pt.cases_per_lane;
    }
    if (num_registers == 16 && t.cases_per_lane > 4) t.cases_per_lane = 4;
    if (t.waves_per_block <= 0 || t.waves_per_block > 4) t.waves_per_block = 4;
    if (t.workers_per_cu <= 0) t.workers_per_cu = 32;
    if (t.chunk <= 0) t.chunk = kMaxQueueGrab;
    if (t.chunk > kMaxQueueGrab) t.chunk = kMaxQueueGrab;
    if (t.program_staging != kStageScalar) t.program_staging = kStageVector;
    return t;
}

template <int NR, int K, int EXT>
static hipError_t launch_batched_t(const DevData& ds, const BatchTuning& t, int n_cu,
                                   const BatchScratch& s, const uint32_t* words,
                                   const int64_t* offs, const int64_t* eff_counts, int64_t cap,
                                   int64_t n, float* fitness, double* sse, hipStream_t stream) {
    const int wpb = t.waves_per_block;
    const int64_t waves = (int64_t)n_cu * t.workers_per_cu;
    const unsigned grid = (unsigned)((waves + wpb - 1) / wpb);
    // at least ~4 grabs per wave: a small population in large chunks would
    // leave most waves without work
    int64_t chunk = n / (waves * 4);
    chunk = chunk < 1 ? 1 : (chunk > t.chunk ? t.chunk : chunk);
    if (t.program_staging == kStageScalar)
        hipLaunchKernelGGL((batched_kernel<NR, K, EXT, kStageScalar>), dim3(grid), dim3(wpb * kWave),
                           0, stream, words, offs, s.order, s.queue_head, n, (int)chunk,
                           eff_counts, cap, ds, fitness, sse);
    else
        hipLaunchKernelGGL((batched_kernel<NR, K, EXT, kStageVector>), dim3(grid), dim3(wpb * kWave),
                           0, stream, words, offs, s.order, s.queue_head, n, (int)chunk,
                           eff_counts, cap, ds, fitness, sse);
    return hipGetLastError();
}

hipError_t launch_batch_order(const int64_t* offsets, int64_t n, const BatchScratch& s,
                              hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    if (n > s.max_programs || n > (int64_t)INT32_MAX - (int64_t(1) << 22)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(kBuckets), 0, stream, s.bucket_count,
                       s.bucket_cursor, s.queue_head);
    const unsigned hgrid = (unsigned)(((n + 255) / 256) < 1024 ? ((n + 255) / 256) : 1024);
    hipLaunchKernelGGL(batch_histogram_kernel, dim3(hgrid), dim3(256), 0, stream, offsets, n,
                       s.bucket_count);
    hipLaunchKernelGGL(batch_bucket_start_kernel, dim3(1), dim3(64), 0, stream, s.bucket_count,
                       s.bucket_cursor);
    hipLaunchKernelGGL(batch_scatter_kernel, dim3((unsigned)((n + kScatterTile - 1) / kScatterTile)),
                       dim3(kScatterThreads), 0, stream, offsets, n, s.bucket_cursor, s.order);
    return hipGetLastError();
}

hipError_t launch_batched_eval(const DevData& ds, int num_registers, BatchTuning t, int n_cu,
                               const BatchScratch& s, const uint32_t* words, const int64_t* offs,
                               const int64_t* eff_counts, int64_t cap, int64_t n,
                               float* fitness, double* sse,
                               hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    if (n > s.max_programs || n > (int64_t)INT32_MAX - (int64_t(1) << 22)) return hipErrorInvalidValue;
    t = resolve_batch_tuning(ds, num_registers, t, n_cu);
#define GALOP_CASE(NRV, KV)                                                             \
    if (num_registers == NRV && t.cases_per_lane == KV)                                 \
        return ds.op_set ? launch_batched_t<NRV, KV, 1>(ds, t, n_cu, s, words, offs, eff_counts, cap, n, fitness, sse, stream) \
                         : launch_batched_t<NRV, KV, 0>(ds, t, n_cu, s, words, offs, eff_counts, cap, n, fitness, sse, stream);
    GALOP_CASE(8, 1) GALOP_CASE(8, 2) GALOP_CASE(8, 4) GALOP_CASE(8, 8)
    GALOP_CASE(16, 1) GALOP_CASE(16, 2) GALOP_CASE(16, 4)
#undef GALOP_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_batched(const DevData& ds, int num_registers, BatchTuning t, int n_cu,
                          const BatchScratch& s, const uint32_t* words, const int64_t* offsets,
                          int64_t n, float* fitness, double* sse, hipStream_t stream) {
    hipError_t e = launch_batch_order(offsets, n, s, stream);
    if (e != hipSuccess) return e;
    return launch_batched_eval(ds, num_registers, t, n_cu, s, words, offsets, nullptr, 0, n,
                               fitness, sse,
                               stream);
}

}  // namespace galop
