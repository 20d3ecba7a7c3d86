#pragma once
// First-pass distance scan on gfx950 MFMA (v_mfma_f32_32x32x16_bf16): kernel
// templates, instantiated per tile configuration by kernels_scan_*.hip.
//
//
// A work item is (a tile of up to BN queries, a contiguous range of base rows).
// The workgroup streams the rows in BM-row tiles; per tile it computes the
// BM x BN block of  a' = |x~|^2 - 2 q~.x~  (bf16 operands, FP32 accumulate;
// a' + |q~|^2 is the squared distance of the bf16 vectors) with operands staged
// through an LDS ring, then runs the epilogue of its mode on the accumulators
// in registers:
//   FILTER : emit (a', row) when a' < tau[q] into the query's private region
//            of this item (no global atomics: counters live in LDS)
//   DENSE  : store every a' (threshold estimation on a row sample)
//   TOP1   : FILTER whose per-lane threshold tightens to the certification band
//            of the smallest a' seen so far (k-means assignment)
//   GATHER : FILTER for work items that share queries (IVF lists): emissions go
//            to an LDS buffer per query and reach the query's global buffer
//            with one atomic per (item, query) at the end of the item
//   SYM    : FILTER over the tiles on and below the diagonal of a self-join
//            (rows = queries, folded norms): S^ is symmetric, so a block below
//            the diagonal also emits (S^, query) for each of its rows whose own
//            threshold it passes.  Those transposed emissions are staged in LDS
//            and appended to the row's transposed region with one global atomic
//            each at the end of the item (an atomic's returned value inside the
//            loop would wait for the chunks in flight); past the staging
//            capacity they are appended at once
// Only candidates cross HBM; the BM x BN block never leaves registers.
//
// Layout: WR x WC waves (rows x queries); a wave owns 32*WM rows x 32*WN
// queries, i.e. WM x WN MFMA blocks.  MFMA 32x32 output: lane l holds query
// column l%32 and rows 8*(e/4) + 4*(l/32) + e%4 of register e.
#include "bounds.h"
#include "launch.h"

namespace mann {

namespace {

typedef __attribute__((address_space(3))) void lds_void;

// s_waitcnt vmcnt(N) lgkmcnt(0) (gfx9 encoding; expcnt left open)
template <int N>
__device__ __forceinline__ void wait_vm_lgkm() {
  __builtin_amdgcn_s_waitcnt((N & 15) | (((N >> 4) & 3) << 14) | (7 << 4));
}
// s_waitcnt lgkmcnt(0) only (vmcnt and expcnt left open)
__device__ __forceinline__ void wait_lgkm0() { __builtin_amdgcn_s_waitcnt(0xC07F); }

// LDS accesses of the epilogue, written out.  They never touch the ring
// stages or norm slots that global->LDS loads in flight are filling, but the
// compiler cannot tell, and before each LDS access it would wait for every
// outstanding vector-memory operation: the next tile's chunks and all earlier
// emission stores.  Each statement waits for its own LDS operation.
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned lds_addr(const void* p) {
  return (unsigned)(uintptr_t)(const lds_void*)p;
}
__device__ __forceinline__ int lds_fetch_add(int* p, int v) {
  int r;
  asm volatile("ds_add_rtn_u32 %0, %1, %2\n\ts_waitcnt lgkmcnt(0)"
               : "=v"(r)
               : "v"(lds_addr(p)), "v"(v)
               : "memory");
  return r;
}
__device__ __forceinline__ floatx4 lds_read_f4(const float* p) {
  floatx4 r;
  asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(r) : "v"(lds_addr(p)) : "memory");
  return r;
}
__device__ __forceinline__ void lds_write_u32(void* p, unsigned v) {
  asm volatile("ds_write_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : : "v"(lds_addr(p)), "v"(v) : "memory");
}
__device__ __forceinline__ void lds_write_u2(uint2* p, uint2 v) {
  const u32x2 w = {v.x, v.y};
  asm volatile("ds_write_b64 %0, %1\n\ts_waitcnt lgkmcnt(0)" : : "v"(lds_addr(p)), "v"(w) : "memory");
}
// Workgroup barrier without the release fence of __syncthreads (which would
// drain every outstanding LDS DMA): callers wait for exactly the chunk they
// need first.  The empty asm statements keep the compiler from moving LDS
// accesses across it.
__device__ __forceinline__ void bare_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// XOR swizzle of the 16-byte chunks of LDS row r (SLOTS = 4 or 8 per row); the
// staging writes stay contiguous per wavefront whatever the swizzle.  For
// 64-byte rows the fragment reads (32 consecutive rows, one logical chunk per
// half-wave) are bank-conflict free on gfx950 with the chunk index XORed by
// bits 3-4 of the row (measured with SQ_LDS_BANK_CONFLICT: 0 cycles, against
// half of all LDS cycles for the XOR of bits 1-2 and 3-4).
template <int SLOTS>
__device__ __forceinline__ int swz(int r) {
  if (SLOTS == 4) return (r >> 3) & 3;
  return (r ^ (r >> 3)) & 7;
}

// Exclusion threshold (squared-distance domain) for the smallest norms-form
// distance, given the first-pass value v of an emitted candidate: any centroid
// whose first-pass value is at least the result has an FP32 norms-form distance
// strictly above the candidate's (norms_form_threshold).
__device__ float top1_threshold(float v, float qe, float qr, float qn2, float ex, float rx,
                                float ra, int d) {
  return (float)norms_form_threshold(v, qe, qr, qn2, ex, rx, ra, d) * (1.0f + 1e-6f);
}

// 16-byte global->LDS load per lane with cache-policy bits AUX (gfx950: 2 =
// non-temporal).
template <int AUX>
__device__ __forceinline__ void gl_lds16(const void* g, lds_void* l) {
  __builtin_amdgcn_global_load_lds(g, l, 16, 0, AUX);
}

// Block tile: WR x WC wavefronts, each owning 32*WM rows x 32*WN queries.
template <int WM, int WN, int WR, int WC, int BK, int ST>
struct Tile {
  static constexpr int NT = 64 * WR * WC;               // threads
  static constexpr int BM = 32 * WM * WR;
  static constexpr int BN = 32 * WN * WC;
  // LDS ring depth: chunks s+1 .. s+STAGES-1 are in flight while s is used
  static constexpr int STAGES = ST;
  static constexpr int SLOTS = BK / 8;                  // 16-byte chunks per LDS row
  static constexpr int A_CHUNKS = BM * SLOTS / NT;      // 16B chunks per thread
  static constexpr int B_CHUNKS = BN * SLOTS / NT;
  static constexpr int LROW = BK;
  // |x~|^2 of the tile's rows, staged by the same DMA as chunk 0 of the tile,
  // in a ring of NXB tiles (a slot is rewritten only after a barrier that
  // follows the epilogue which read it, whatever the chunks per tile)
  static constexpr int NXB = 4;
  static_assert(WR * WC >= BM / 64, "one wavefront stages the norms of each 64 rows");
  static constexpr size_t LDS_BYTES =
      (size_t)STAGES * (BM + BN) * LROW * sizeof(bf16_t) + BN * 4 + (size_t)NXB * BM * 4;
  static constexpr int NPC = A_CHUNKS + B_CHUNKS;       // DMA ops per thread per chunk
  // GATHER mode: per query column a small LDS candidate buffer
  static constexpr int LCAP = 16;
  static constexpr size_t LDS_GATHER = LDS_BYTES + (size_t)BN * LCAP * 8;
  // SYM mode: staged transposed emissions (row, value, query) of the item
  static constexpr int SCAP = 4096;
  static constexpr size_t LDS_SYM = LDS_BYTES + 16 + (size_t)SCAP * 12;
};

// One work item.  Values are in the squared-distance domain:
//   S^ = |x~|^2 - 2 q~.x~ + |q~|^2   (fma, then the query's norm added)
// Loop nest: row tiles outside, K chunks inside, so the accumulators stay in
// place across the K loop (no copies at a loop-carried merge).  Chunks of the
// flattened (tile, K chunk) sequence stream through an LDS ring of STAGES
// stages by direct global->LDS loads, STAGES-1 chunks ahead; within a chunk the
// fragment reads of the next K step are in flight while the current one is
// multiplied out, and the first step's reads of chunk s+1 follow the barrier
// that ends chunk s.
template <int WM, int WN, int WR, int WC, int BK, int ST, int MODE>
__device__ __forceinline__ void scan_item(const ScanArgs& a, int qoff, int nqi, int row_begin,
                                          int row_end, int region, unsigned char* smem) {
  using T = Tile<WM, WN, WR, WC, BK, ST>;
  constexpr int BM = T::BM, BN = T::BN, LROW = T::LROW, SL = T::SLOTS, STAGES = T::STAGES;
  constexpr int NT = T::NT;
  bf16_t* As = (bf16_t*)smem;                   // [STAGES][BM][LROW]
  bf16_t* Bs = As + STAGES * BM * LROW;         // [STAGES][BN][LROW]
  int* lcnt = (int*)(Bs + STAGES * BN * LROW);  // [BN] emitted count per query column
  float* nxs = (float*)(lcnt + BN);             // [NXB][BM] row norms of the tiles
  uint2* lbuf = (uint2*)(nxs + T::NXB * BM);    // [BN][LCAP] (GATHER mode)
  // SYM mode: [0] staged count, [1] overflow flag, then rows, values, queries
  int* scnt = (int*)(nxs + T::NXB * BM);
  unsigned* srow = (unsigned*)(scnt + 4);
  float* sval = (float*)(srow + T::SCAP);
  unsigned* scol = (unsigned*)(sval + T::SCAP);
  constexpr bool SYM = MODE == SCAN_SYM;
  // a SYM item's query tile and whether a tile of rows lies below it
  const int sym_ctile = qoff / BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave % WR, wc = wave / WR;
  const int kp = a.kp, nk = kp / BK;
  const int ntiles = (row_end - row_begin + BM - 1) / BM;
  if (ntiles <= 0 || nqi <= 0) return;

  const int* qidx = a.qidx;
  // Staging: thread chunk c = tid + NT p lands at LDS row c/SL, slot c%SL (a
  // contiguous 16-byte write); it carries logical K chunk slot ^ swz(row).
  // byte offsets: within a row tile 32 bits suffice (a uniform base plus a
  // 32-bit lane offset, the scalar-base form of the DMA); query rows can lie
  // anywhere in a large operand
  unsigned offA[T::A_CHUNKS];
  int64_t offB[T::B_CHUNKS];
#pragma unroll
  for (int p = 0; p < T::A_CHUNKS; ++p) {
    const int c = tid + NT * p, r = c / SL;
    offA[p] = (unsigned)(r * kp + (((c % SL) ^ swz<SL>(r)) * 8)) * 2u;
  }
#pragma unroll
  for (int p = 0; p < T::B_CHUNKS; ++p) {
    const int c = tid + NT * p, r = c / SL;
    int brow;
    if (a.b_by_slot)
      brow = qoff + (r < nqi ? r : 0);
    else
      brow = r < nqi ? (qidx ? qidx[qoff + r] : qoff + r) : (qidx ? qidx[qoff] : qoff);
    offB[p] = ((int64_t)brow * kp + (((c % SL) ^ swz<SL>(r)) * 8)) * 2;
  }
  int gqc[WN];
  float tau_l[WN], qo[WN];
  // TOP1: the query's rounding norms for the certification band, loaded here
  // (a global load in the epilogue would wait for every store before it)
  float t1e[MODE == SCAN_TOP1 ? WN : 1], t1r[MODE == SCAN_TOP1 ? WN : 1];
  float t1n[MODE == SCAN_TOP1 ? WN : 1];
#pragma unroll
  for (int bn = 0; bn < WN; ++bn) {
    int c = wc * 32 * WN + bn * 32 + (lane & 31);
    int gq = c < nqi ? (qidx ? qidx[qoff + c] : qoff + c) : -1;
    gqc[bn] = gq;
    tau_l[bn] = (MODE != SCAN_DENSE && gq >= 0) ? a.tau[gq] : -INFINITY;
    float o = 0.0f;
    if (gq >= 0) o = a.col_norm ? a.col_norm[qoff + c] : (a.qnorm ? a.qnorm[gq] : 0.0f);
    qo[bn] = o;
    if (MODE == SCAN_TOP1) {
      t1e[bn] = gq >= 0 ? a.t1_qe[gq] : 0.0f;
      t1r[bn] = gq >= 0 ? a.t1_qr[gq] : 0.0f;
      t1n[bn] = gq >= 0 ? a.t1_qn2[gq] : 0.0f;
    }
  }

  // chunk (t, kc) -> LDS stage st by direct global->LDS loads: wave w's
  // instruction p fills 1 KiB at chunk slots [NT p + 64w, +64) (lane order)
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  // row norms: wave w stages rows (w mod BM/64)*64 + lane of the tile (one
  // 4-byte DMA per lane; waves beyond BM/64 repeat the same values).  Like the
  // row operand, the norms of a whole tile are read: every norm array has at
  // least 256 padding rows past its last row (+inf), and the epilogue masks
  // rows at or past row_end.
  // (A buffer load: scalar row offset + 32-bit lane offset, so no 64-bit
  // per-lane address is held across the loop.)
  const int nx_slot = (wv % (BM / 64)) * 64;
  const int cpol = __builtin_amdgcn_readfirstlane(a.cpol);
  // SYM stages each row's own threshold instead of its norm (folded norms need
  // none): its caller's threshold array holds at least 256 entries past the
  // last row (read with the rest of the last tile, masked by the row test)
  const __amdgpu_buffer_rsrc_t nx_rsrc =
      __builtin_amdgcn_make_buffer_rsrc(SYM ? (void*)a.tau : (void*)a.xn, 0, 0xffffffffu,
                                        0x00020000);
  auto dma = [&](int t, int kc, int st) {
    if (kc == 0 && (!a.fold || SYM))
      __builtin_amdgcn_raw_ptr_buffer_load_lds(nx_rsrc,
                                               (lds_void*)(nxs + (t % T::NXB) * BM + nx_slot), 4,
                                               lane * 4,
                                               (int)((unsigned)(row_begin + t * BM + nx_slot) * 4u),
                                               0, 0);
    const bf16_t* baseA = a.xb + (int64_t)(row_begin + t * BM) * kp + kc * BK;
    const bf16_t* baseB = a.qb + kc * BK;
    if (cpol & 1) {
#pragma unroll
      for (int p = 0; p < T::A_CHUNKS; ++p)
        gl_lds16<2>((const char*)baseA + offA[p],
                    (lds_void*)(As + st * BM * LROW + (NT * p + 64 * wv) * 8));
    } else {
#pragma unroll
      for (int p = 0; p < T::A_CHUNKS; ++p)
        gl_lds16<0>((const char*)baseA + offA[p],
                    (lds_void*)(As + st * BM * LROW + (NT * p + 64 * wv) * 8));
    }
    if (cpol & 2) {
#pragma unroll
      for (int p = 0; p < T::B_CHUNKS; ++p)
        gl_lds16<2>((const char*)baseB + offB[p],
                    (lds_void*)(Bs + st * BN * LROW + (NT * p + 64 * wv) * 8));
    } else {
#pragma unroll
      for (int p = 0; p < T::B_CHUNKS; ++p)
        gl_lds16<0>((const char*)baseB + offB[p],
                    (lds_void*)(Bs + st * BN * LROW + (NT * p + 64 * wv) * 8));
    }
  };
  // wait until at most `later` chunks issued after the one needed next are in
  // flight (a chunk 0 carries one more op, the row norms: counting NPC per
  // chunk can only wait longer, never shorter, since loads retire in order)
  auto wait_chunks = [&](int later) {
    constexpr int N = T::NPC;
    static_assert(STAGES >= 3 && STAGES <= 5 && (STAGES - 2) * N <= 63, "vmcnt range");
    if (later <= 0) wait_vm_lgkm<0>();
    else if (later == 1 || STAGES == 3) wait_vm_lgkm<N>();
    else if (later == 2 || STAGES == 4) wait_vm_lgkm<2 * N>();
    else wait_vm_lgkm<3 * N>();
  };

  floatx16 acc[WM][WN];

  // Fragments of one 16-wide K step, double-buffered in registers: the reads
  // of step ks+1 are in flight while step ks is multiplied out.
  constexpr int NKS = BK / 16;
  static_assert(NKS % 2 == 0, "K steps per chunk must be even");
  bf16x8 F[2][WM + WN];
  auto rd = [&](int b, int buf, int ks) {
    const bf16_t* Ab = As + buf * BM * LROW;
    const bf16_t* Bb = Bs + buf * BN * LROW;
    const int kc = ks * 2 + (lane >> 5);  // logical 16-byte chunk of this lane
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const int r = wr * 32 * WM + i * 32 + (lane & 31);
      F[b][i] = *(const bf16x8*)(Ab + r * LROW + ((kc ^ swz<SL>(r)) * 8));
    }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int r = wc * 32 * WN + j * 32 + (lane & 31);
      F[b][WM + j] = *(const bf16x8*)(Bb + r * LROW + ((kc ^ swz<SL>(r)) * 8));
    }
  };
  // zero: first step of a tile, accumulate onto 0 (no register clearing)
  auto mm = [&](int b, bool zero) {
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(F[b][i], F[b][WM + j],
                                                             zero ? floatx16{} : acc[i][j], 0,
                                                             0, 0);
    __builtin_amdgcn_s_setprio(0);
  };

  // Epilogue of one BM-row tile: per MFMA block, S^ from the accumulators,
  // threshold mask, then the rare emissions.
  // Epilogue of one BM-row tile.  FOLD: the accumulators are S^ already
  // (folded norms); FULL: no row of the tile lies at or past row_end.
  auto epilogue_t = [&](int t, auto FOLDC, auto FULLC) {
    constexpr bool FOLD = decltype(FOLDC)::value, FULL = decltype(FULLC)::value;
    // lane-derived addresses are recomputed from an opaque copy of the lane
    // id here, rather than hoisted out of the tile loop into registers that
    // the K loop needs
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int r0 = row_begin + t * BM;
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const int rb0 = r0 + wr * 32 * WM + i * 32 + 4 * (ln >> 5);
      const float* nxt = nxs + (t % T::NXB) * BM + (rb0 - r0);
      // non-folded: |x~|^2 of this lane's 16 rows (+inf past row_end), from
      // LDS; folded: 0, or +inf past row_end
      float nx[16];
#pragma unroll
      for (int e4 = 0; e4 < 4; ++e4) {
        floatx4 v4 = {0.0f, 0.0f, 0.0f, 0.0f};
        if (!FOLD) v4 = lds_read_f4(nxt + 8 * e4);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int row = rb0 + 8 * e4 + c;
          nx[4 * e4 + c] = row < row_end ? v4[c] : INFINITY;
        }
      }
      // first-pass value without the query's norm (folded: with it)
      auto val = [&](int j, int e) -> float {
        if (FOLD) return FULL ? acc[i][j][e] : acc[i][j][e] + nx[e];
        return __builtin_fmaf(-2.0f, acc[i][j][e], nx[e]);
      };
      // fast test of every block of this row group first: one FMA per value
      // and a balanced min tree per block (independent chains); fl(min_e a_e
      // + qo) = min_e fl(a_e + qo) because rounding is monotone, so "no value
      // of this block passes" is decided exactly
      // FILTER with folded norms: the values are S^ themselves, so one compare
      // per value serves the block test, the emission count and the store masks
      constexpr bool CMP = FOLD && (MODE == SCAN_FILTER || MODE == SCAN_SYM);
      // SYM below the diagonal: the thresholds of this lane's 16 rows, staged
      // with the tile (rows at or past row_end are masked by their values)
      float trow[SYM ? 16 : 1];
      const bool offdiag = SYM && (r0 / BM) > sym_ctile;
      if (SYM && offdiag) {
#pragma unroll
        for (int e4 = 0; e4 < 4; ++e4) {
          const floatx4 v4 = lds_read_f4(nxs + (t % T::NXB) * BM + (rb0 - r0) + 8 * e4);
#pragma unroll
          for (int c = 0; c < 4; ++c) trow[SYM ? 4 * e4 + c : 0] = v4[c];
        }
      }
      bool pass[WN];
      if (MODE != SCAN_DENSE && !CMP) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < WN; ++j) {
          float m4[4];
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            m4[g] = fminf(fminf(val(j, 4 * g), val(j, 4 * g + 1)),
                          fminf(val(j, 4 * g + 2), val(j, 4 * g + 3)));
          }
          const float mn = fminf(fminf(m4[0], m4[1]), fminf(m4[2], m4[3]));
          pass[j] = (FOLD ? mn : mn + qo[j]) < tau_l[j];
        }
      }
#pragma unroll
      for (int j = 0; j < WN; ++j) {
        if (SYM && offdiag) {
          // transposed emissions: query = this lane's row, candidate = its column
          bool tp[16];
          uint64_t tm[16];
          int tot = 0;
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            tp[e] = gqc[j] >= 0 && val(j, e) < trow[SYM ? e : 0];
            tm[e] = __builtin_amdgcn_ballot_w64(tp[e]);
            tot += __popcll(tm[e]);
          }
          if (tot) {  // wave-uniform
            int base = 0;
            if (ln == 0) base = lds_fetch_add(scnt, tot);
            base = __shfl(base, 0);
            const uint64_t lt = (1ull << ln) - 1ull;
            int gq = gqc[j];
            asm volatile("" : "+v"(gq));
#pragma unroll
            for (int e = 0; e < 16; ++e) {
              if (tp[e]) {
                const int pos = base + __popcll(tm[e] & lt);
                const unsigned row = (unsigned)(rb0 + 8 * (e >> 2) + (e & 3));
                if (pos < T::SCAP) {
                  lds_write_u32(&srow[pos], row);
                  lds_write_u32(&sval[pos], __float_as_uint(val(j, e)));
                  lds_write_u32(&scol[pos], (unsigned)gq);
                } else {
                  // staging full: straight to the row's region, so that a
                  // region's count never exceeds the entries written to it
                  // (before the capacity): the selection reads that many
                  const int rs = a.rstride ? a.rstride : a.parts;
                  const int64_t slot =
                      (int64_t)row * rs + a.parts + (qoff / a.rows_per_part);
                  const int p2 = atomicAdd(a.cnt + slot, 1);
                  if (p2 < a.cap) {
                    uint2 cv;
                    cv.x = __float_as_uint(val(j, e));
                    cv.y = (unsigned)gq;
                    *(uint2*)(a.cand + (slot * a.cap + p2) * 2) = cv;
                  }
                }
              }
              base += __popcll(tm[e]);
            }
          }
        }
        if (MODE == SCAN_DENSE) {
          // (blocks are counted from the list's first row with dense_rel, else
          // from row 0 of the scanned rows)
          const int gbase = a.dense_rel ? region : 0;
          if (a.dense_grp > 0 && row_end - gbase >= a.dense_grp) {
            // minima of the lane's four runs of 4 consecutive rows (rows past
            // row_end are +inf), one 16-byte store per lane and 32-row block
            const int b0 = rb0 - 4 * (ln >> 5);  // the block's first row
            floatx4 v4;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              float mn = fminf(fminf(val(j, 4 * u), val(j, 4 * u + 1)),
                               fminf(val(j, 4 * u + 2), val(j, 4 * u + 3)));
              v4[u] = FOLD ? mn : mn + qo[j];
            }
            if (gqc[j] >= 0 && b0 < row_end) {
              const int blk = (b0 - gbase) >> 5;
              *(floatx4*)(a.dense + (int64_t)gqc[j] * a.ldd + blk * 8 + (ln >> 5) * 4) = v4;
            }
            continue;
          }
          if (a.dense_bmin) {
            // the minimum of each 32-row block: 16 of its values here, 16 in
            // lane ^ 32 (fl(min a + qo) = min fl(a + qo): monotone rounding)
            float mn = val(j, 0);
#pragma unroll
            for (int e = 1; e < 16; ++e) mn = fminf(mn, val(j, e));
            if (!FOLD) mn += qo[j];
            mn = fminf(mn, __shfl_xor(mn, 32));
            const int blk = (rb0 - 4 * (ln >> 5)) >> 5;
            if (ln < 32 && gqc[j] >= 0 && (blk << 5) < row_end)
              a.dense[(int64_t)gqc[j] * a.ldd + blk] = mn;
            continue;
          }
          if (gqc[j] >= 0) {
            // the lane's 16 rows are four runs of 4 consecutive rows: with
            // every row inside, four 16-byte stores (the caller keeps ldd, the
            // base and the runs' offsets multiples of 4 floats when dvec)
            if (!a.dense_t && a.dvec && rb0 + 27 < row_end) {
              float* dst =
                  a.dense + (int64_t)gqc[j] * a.ldd + (a.dense_rel ? rb0 - region : rb0);
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                floatx4 v4;
#pragma unroll
                for (int c2 = 0; c2 < 4; ++c2)
                  v4[c2] = FOLD ? val(j, 4 * u + c2) : val(j, 4 * u + c2) + qo[j];
                *(floatx4*)(dst + 8 * u) = v4;
              }
              continue;
            }
#pragma unroll
            for (int e = 0; e < 16; ++e) {
              const int row = rb0 + 8 * (e >> 2) + (e & 3);
              const float v = FOLD ? val(j, e) : val(j, e) + qo[j];
              if (row < row_end) {
                if (a.dense_t)
                  a.dense[(int64_t)row * a.ldd + gqc[j]] = v;
                else
                  a.dense[(int64_t)gqc[j] * a.ldd + (a.dense_rel ? row - region : row)] = v;
              }
            }
          }
          continue;
        }
        if (CMP) {
          bool p[16];
          uint64_t any = 0;
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            p[e] = val(j, e) < tau_l[j];
            any |= __builtin_amdgcn_ballot_w64(p[e]);
          }
          if (any == 0) continue;  // wave-uniform
          int cnt = 0;
#pragma unroll
          for (int e = 0; e < 16; ++e) cnt += p[e] ? 1 : 0;
          int gq = gqc[j];
          asm volatile("" : "+v"(gq));
          const int qc = wc * 32 * WN + j * 32 + (ln & 31);
          int pos = lds_fetch_add(&lcnt[qc], cnt);
          const int rs = a.rstride ? a.rstride : a.parts;
          unsigned* reg = a.cand + ((int64_t)gq * rs + region) * (int64_t)a.cap * 2;
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            if (!p[e]) continue;
            if (pos < a.cap) {
              uint2 cv;
              cv.x = __float_as_uint(val(j, e));
              cv.y = (unsigned)(rb0 + 8 * (e >> 2) + (e & 3));
              *(uint2*)(reg + 2 * pos) = cv;
            }
            ++pos;
          }
          continue;
        }
        if (!pass[j]) continue;
        // the rare emissions: S^ of the block again (same operations), mask
        float w[16];
        unsigned m = 0;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          w[e] = FOLD ? val(j, e) : val(j, e) + qo[j];
          m |= (w[e] < tau_l[j] ? 1u : 0u) << e;
        }
        int gq = gqc[j];
        asm volatile("" : "+v"(gq));
        const int qc = wc * 32 * WN + j * 32 + (ln & 31);  // query column in the tile
        int pos = lds_fetch_add(&lcnt[qc], __builtin_popcount(m));
        unsigned* reg = a.cand + ((int64_t)gq * a.parts + region) * (int64_t)a.cap * 2;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          if (!((m >> e) & 1u)) continue;
          const int row = rb0 + 8 * (e >> 2) + (e & 3);
          uint2 cv;
          cv.x = __float_as_uint(w[e]);
          cv.y = (unsigned)row;
          if (MODE == SCAN_GATHER) {
            if (pos < T::LCAP) {
              lds_write_u2(&lbuf[qc * T::LCAP + pos], cv);
            } else {  // spill straight to the query's global buffer
              int gp = atomicAdd(a.cnt + gq, 1);
              if (gp < a.cap) *(uint2*)(a.cand + ((int64_t)gq * a.cap + gp) * 2) = cv;
            }
          } else if (pos < a.cap) {
            *(uint2*)(reg + 2 * pos) = cv;
          }
          ++pos;
          if (MODE == SCAN_TOP1) {
            float th = top1_threshold(w[e], t1e[MODE == SCAN_TOP1 ? j : 0],
                                      t1r[MODE == SCAN_TOP1 ? j : 0],
                                      t1n[MODE == SCAN_TOP1 ? j : 0], a.t1_ex, a.t1_rx,
                                      a.t1_ra, a.d);
            tau_l[j] = fminf(tau_l[j], th);
          }
        }
      }
    }
  };
  auto epilogue = [&](int t) {
    using T1 = std::integral_constant<bool, true>;
    using T0 = std::integral_constant<bool, false>;
    if (MODE != SCAN_TOP1 && MODE != SCAN_GATHER && (a.fold || SYM)) {
      if (row_begin + (t + 1) * BM <= row_end) epilogue_t(t, T1{}, T1{});
      else epilogue_t(t, T1{}, T0{});
    } else {
      epilogue_t(t, T0{}, T0{});
    }
  };

  __syncthreads();  // LDS may still be in use by a previous item of this workgroup
  for (int c = tid; c < BN; c += NT) lcnt[c] = 0;
  if (SYM && tid == 0) {
    scnt[0] = 0;
    scnt[1] = 0;
  }
  const int S = ntiles * nk;
  constexpr int D = STAGES - 1;  // prefetch distance in chunks
  // issue chunks 0 .. D-1
  int ti = 0, ki = 0, issued = 0;
  for (int i = 0; i < D && i < S; ++i) {
    dma(ti, ki, i);
    ++issued;
    if (++ki == nk) {
      ki = 0;
      ++ti;
    }
  }
  wait_chunks(issued - 1);
  __syncthreads();
  int s = 0, st0 = 0, sti = D;  // chunk s in stage st0; next issue goes to stage sti
  // One chunk: K steps 0 .. NKS-1 out of stage st0.  The reads of the first
  // step of the next chunk follow the barrier and overlap the last step's
  // multiply-adds; for the last chunk of a tile they wait until after the
  // epilogue, which needs the registers (after the last chunk of the item they
  // read a stale stage, unused).
  auto chunk = [&](bool first, bool last) __attribute__((always_inline)) {
    if (issued < S) {  // chunk s + D into the stage chunk s-1 used
      dma(ti, ki, sti);
      ++issued;
      if (++ki == nk) {
        ki = 0;
        ++ti;
      }
      sti = sti == STAGES - 1 ? 0 : sti + 1;
    }
#pragma unroll
    for (int ks = 0; ks + 1 < NKS; ++ks) {
      // step ks's fragments (read a whole K step earlier) have landed; waiting
      // here, before the next reads are issued, keeps the multiply-adds below
      // from waiting on those
      wait_lgkm0();
      rd((ks + 1) & 1, st0, ks + 1);
      mm(ks & 1, first && ks == 0);
    }
    // chunk s+1 must have landed; chunks s+2 .. issued-1 may stay in flight
    wait_chunks(issued - (s + 2));
    bare_barrier();
    wait_lgkm0();  // no-op (waited above); tells the compiler step NKS-1 has landed
    st0 = st0 == STAGES - 1 ? 0 : st0 + 1;
    if (!last) rd(0, st0, 0);
    mm((NKS - 1) & 1, false);
    ++s;
  };
  rd(0, st0, 0);
  for (int t = 0; t < ntiles; ++t) {
    if (nk == 1) {
      chunk(true, true);
    } else {
      chunk(true, false);
      for (int kc = 1; kc + 1 < nk; ++kc) chunk(false, false);
      chunk(false, true);
    }
    epilogue(t);
    rd(0, st0, 0);
  }

  if (MODE == SCAN_GATHER) {
    __syncthreads();
    for (int c = tid; c < nqi; c += NT) {
      const int n = min(lcnt[c], T::LCAP);
      if (n == 0) continue;
      const int gq = qidx ? qidx[qoff + c] : qoff + c;
      const int base = atomicAdd(a.cnt + gq, n);
      unsigned* dst = a.cand + (int64_t)gq * a.cap * 2;
      for (int i = 0; i < n && base + i < a.cap; ++i)
        *(uint2*)(dst + 2 * (base + i)) = lbuf[c * T::LCAP + i];
    }
  } else if (MODE != SCAN_DENSE) {
    __syncthreads();
    const int rs = a.rstride ? a.rstride : a.parts;
    for (int c = tid; c < nqi; c += NT) {
      int gq = qidx ? qidx[qoff + c] : qoff + c;
      a.cnt[(int64_t)gq * rs + region] = lcnt[c];
    }
    if (SYM) {
      // the transposed emissions go to each row's region for the partition of
      // this item's query tile (several items append to it: one atomic each)
      const int creg = a.parts + (qoff / a.rows_per_part);
      const int ns = min(scnt[0], T::SCAP);
      for (int i = tid; i < ns; i += NT) {
        const int64_t slot = (int64_t)srow[i] * rs + creg;
        const int pos = atomicAdd(a.cnt + slot, 1);
        if (pos < a.cap) {
          uint2 cv;
          cv.x = __float_as_uint(sval[i]);
          cv.y = scol[i];
          *(uint2*)(a.cand + (slot * a.cap + pos) * 2) = cv;
        }
      }
    }
  }
}

// Wave tiles of at most 128 x 64 keep two waves per SIMD.
template <int WM, int WN, int WR, int WC, int BK, int ST, int MODE, int REFINE>
__global__ __launch_bounds__(64 * WR * WC)
__attribute__((amdgpu_waves_per_eu(WM * WN <= 8 ? 2 : 1))) void scan_grid_kernel(ScanArgs a) {
  using T = Tile<WM, WN, WR, WC, BK, ST>;
  extern __shared__ __align__(16) unsigned char smem[];
  int nq = REFINE ? *a.nq_dev : a.nq;
  int bid = blockIdx.x;
  const int G = gridDim.x;
  // XCD-aware order: consecutive logical ids share an XCD (and its L2)
  if (a.xcd_map && (G & 7) == 0) bid = (bid & 7) * (G >> 3) + (bid >> 3);
  int qtile = 0, part = 0;
  if (MODE == SCAN_SYM) {
    // (decoded below)
  } else if (a.xcd_map == 2) {  // part-major: an XCD streams one row partition
    part = bid / a.qtiles;
    qtile = bid - part * a.qtiles;
  } else if (a.xcd_map == 3) {
    // bands of 8 row partitions: the ~64 workgroups an XCD runs at once cover
    // 8 partitions x 8 query tiles, so each row chunk and each query chunk
    // they load is shared by 8 of them through the XCD's L2
    const int SPB = a.xcd_band > 0 ? a.xcd_band : 8;
    const int g = bid / (SPB * a.qtiles);
    const int p0 = g * SPB;
    const int bp = min(SPB, a.parts - p0);
    const int idx = bid - g * SPB * a.qtiles;
    part = p0 + idx % bp;
    qtile = idx / bp;
  } else {
    qtile = bid / a.parts;
    part = bid - qtile * a.parts;
  }
  if (MODE == SCAN_SYM) {
    // partition-major items (query tile C, row partition p) with C's first
    // row below the partition's end; the rows start at C's own tile
    int rem = bid;
    part = -1;
    for (int p = 0; p < a.parts; ++p) {
      const int cnt_p = min(a.qtiles, ((p + 1) * a.rows_per_part + T::BN - 1) / T::BN);
      if (rem < cnt_p) {
        part = p;
        qtile = rem;
        break;
      }
      rem -= cnt_p;
    }
    if (part < 0) return;
  }
  int qoff = qtile * T::BN;
  if (qoff >= nq || part >= a.parts) return;
  int nqi = min(T::BN, nq - qoff);
  int rb = part * a.rows_per_part;
  if (MODE == SCAN_SYM) rb = max(rb, qoff);
  int re = min(a.n_rows, part * a.rows_per_part + a.rows_per_part);
  scan_item<WM, WN, WR, WC, BK, ST, MODE>(a, qoff, nqi, rb, re, part, smem);
}

template <int WM, int WN, int WR, int WC, int BK, int ST, int MODE>
__global__ __launch_bounds__(64 * WR * WC)
__attribute__((amdgpu_waves_per_eu(WM * WN <= 8 ? 2 : 1))) void scan_items_kernel(ScanArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_item;
  const int n_items = *a.n_items_dev;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(a.item_counter, 1);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= n_items) break;
    const ScanItem w = a.items[it];
    scan_item<WM, WN, WR, WC, BK, ST, MODE>(a, w.qoff, w.nq, w.row_begin, w.row_end, w.region, smem);
  }
}

template <class K>
void set_lds(K kern, size_t lds) {
  MANN_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
}

template <int WM, int WN, int WR, int WC, int BK, int ST, int MODE>
void launch_t(const ScanArgs& a, int grid, int kind, hipStream_t st) {
  using T = Tile<WM, WN, WR, WC, BK, ST>;
  const size_t lds =
      MODE == SCAN_GATHER ? T::LDS_GATHER : (MODE == SCAN_SYM ? T::LDS_SYM : T::LDS_BYTES);
  MANN_REQUIRE(lds <= (size_t)160 * 1024, ERR_ARGUMENT,
               "this tile configuration needs " + std::to_string(lds) +
                   " bytes of LDS in this scan mode, more than a gfx950 CU holds");
  static bool attr_set = false;  // dynamic LDS above 64 KiB must be opted into
  if (!attr_set) {
    set_lds(scan_items_kernel<WM, WN, WR, WC, BK, ST, MODE>, lds);
    set_lds(scan_grid_kernel<WM, WN, WR, WC, BK, ST, MODE, 0>, lds);
    set_lds(scan_grid_kernel<WM, WN, WR, WC, BK, ST, MODE, 1>, lds);
    attr_set = true;
  }
  const dim3 block(T::NT);
  if (kind == SCAN_KIND_ITEMS)
    hipLaunchKernelGGL((scan_items_kernel<WM, WN, WR, WC, BK, ST, MODE>), dim3(grid), block, lds, st,
                       a);
  else if (kind == SCAN_KIND_REFINE)
    hipLaunchKernelGGL((scan_grid_kernel<WM, WN, WR, WC, BK, ST, MODE, 1>), dim3(grid), block, lds,
                       st, a);
  else
    hipLaunchKernelGGL((scan_grid_kernel<WM, WN, WR, WC, BK, ST, MODE, 0>), dim3(grid), block, lds,
                       st, a);
}

template <int WM, int WN, int WR, int WC, int ST>
void launch_cfg(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st) {
  switch (mode) {
    case SCAN_FILTER: launch_t<WM, WN, WR, WC, 32, ST, SCAN_FILTER>(a, grid, kind, st); break;
    case SCAN_DENSE: launch_t<WM, WN, WR, WC, 32, ST, SCAN_DENSE>(a, grid, kind, st); break;
    case SCAN_TOP1: launch_t<WM, WN, WR, WC, 32, ST, SCAN_TOP1>(a, grid, kind, st); break;
    case SCAN_GATHER: launch_t<WM, WN, WR, WC, 32, ST, SCAN_GATHER>(a, grid, kind, st); break;
    case SCAN_SYM:
      if constexpr (WM * WR == WN * WC)
        launch_t<WM, WN, WR, WC, 32, ST, SCAN_SYM>(a, grid, kind, st);
      else
        throw Error{ERR_ASSERTION, "the symmetric join needs square tiles"};
      break;
    default: break;
  }
}

}  // namespace

}  // namespace mann
