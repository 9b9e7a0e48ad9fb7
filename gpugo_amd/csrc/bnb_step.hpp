// bnb_step.hpp -- the objective-independent kernels of one branch-and-bound
// step: select (the K smallest boxes, stop tests, new parent records),
// compact (list prune) and the root initialisation.  Included by
// libgpugo.so and, as text, by every objective module (gpugo_amd.emit), so a
// step's three launches come from one code object (switching code objects
// between dependent launches costs several microseconds each).
//
// Reference correspondence (paths relative to GPUGO_1.1/):
//   select_body   list prune + argmin pop + completeRegion + cycle update
//                 (1-Ackley/ackley50.py:1154-1220, :1270-1370, :820-948),
//                 generalised to the K smallest boxes per step
//   compact_body  constructFlagArr + mask compaction (:768-810, :1217-1242)
#pragma once

namespace gpugo {

// ---------------------------------------------------------------------------
// block helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned wave_incl_scan(unsigned v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    unsigned t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// exclusive scan over the block; sh holds >= nwaves+1 entries
__device__ unsigned block_excl_scan(unsigned v, unsigned *sh, unsigned *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned inc = wave_incl_scan(v);
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  if (w == 0) {
    unsigned x = lane < nw ? sh[lane] : 0u;
    unsigned xi = wave_incl_scan(x);
    if (lane < nw) sh[lane] = xi - x;
    if (lane == nw - 1) sh[nw] = xi;
  }
  __syncthreads();
  unsigned r = sh[w] + inc - v;
  *total = sh[nw];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// decode (round-to-nearest, one fma per bound, top child ends at U)
// 1-Ackley/ackley50.py:565-593 (device) and :820-948 (host)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void decode_box(const double *plo, const double *phi, const double *pw, int win, int d, int wid, int c,
                           unsigned pos, double *lo, double *hi) {
  for (int j = 0; j < d; ++j) {
    lo[j] = plo[j];
    hi[j] = phi[j];
  }
  unsigned r = pos;
  for (int k = 0; k < c; ++k) {
    int j = win + k;
    if (j >= d) j -= d;
    unsigned q = r / (unsigned)wid, dg = r - q * (unsigned)wid;
    r = q;
    lo[j] = __builtin_fma(pw[j], (double)dg, plo[j]);
    hi[j] = (dg + 1 == (unsigned)wid) ? phi[j] : __builtin_fma(pw[j], (double)(dg + 1), plo[j]);
  }
}

// Coordinate j of child `pos` of a parent record: the same arithmetic as
// decode_box for one coordinate (window coordinates (win + k) mod d, k < c,
// take base-wid digit k of pos).
__device__ __forceinline__ void child_coord(const double *plo, const double *phi, const double *pw, int win, int d,
                                            int wid, int c, unsigned pos, int j, double &lo, double &hi) {
  int k = j - win;
  if (k < 0) k += d;
  if (k < c) {
    unsigned r = pos;
    for (int t = 0; t < k; ++t) r /= (unsigned)wid;
    const unsigned dg = r % (unsigned)wid;
    lo = __builtin_fma(pw[j], (double)dg, plo[j]);
    hi = (dg + 1 == (unsigned)wid) ? phi[j] : __builtin_fma(pw[j], (double)(dg + 1), plo[j]);
  } else {
    lo = plo[j];
    hi = phi[j];
  }
}

// ---------------------------------------------------------------------------
// select: the K smallest (lb, parent, pos) keys of the live queue, exact,
// in one workgroup (radix select on 128-bit keys, 12-bit digits), then the
// stop tests and materialisation of the selected boxes as new parents.
// ---------------------------------------------------------------------------
#define SEL_T 1024
#define SEL_BINS 4096
#define SEL_FINAL 1024
#define SEL_KMAX 1024
#define SEL_U 8
#define GPUGO_DMAX_SELECT (1 << 17)

// Priority band of a queue (a batch slot's, or a single solve's with more
// than one box per step): the compaction copies every kept entry with
// lb < g into a side list (at most cap entries, count in n; more than cap
// leaves it unusable), and the select takes the K smallest from it instead
// of passing over the whole queue whenever it holds at least K entries:
// every entry outside it has lb >= g, above all of them.  The select of a
// step resets n and sets g for its compaction from the histogram of a full
// pass (about BAND_TARGET entries below g); a queue of at most BAND_MIN
// entries keeps no band (g = -inf) and is passed over whole.
struct BandState {
  double g;                 // threshold (-inf: no band)
  unsigned long long n;     // entries appended by the last compaction
};
struct BandView {
  const double *lb;
  const unsigned *par, *pos, *idx;  // idx: the entry's index in the compacted queue
  BandState *st;                    // null: no band
  long long cap;
};
#define BAND_TARGET 2048
#define BAND_MIN 32768

typedef unsigned __int128 u128;

struct SelectArgs {
  DevState *st;
  double *q_lb;
  unsigned *q_par, *q_pos;
  double *plo, *phi, *pw;
  int *pwin;
  double *out_lo, *out_hi, *out_lb;
  long long out_cap, p_cap;
  int d, wid, c, k_select;
  double tol_f, tol_x;
  long long max_iters;   // > 0: stop with GPUGO_ST_MAX_ITERS on the device (batch slots)
  int pstride;           // doubles per parent record (d, or the batch's dmax)
  int ext_n;             // doubles of per-record payload in pext (moved with the records)
  double *pext;          // [p_cap][ext_n] (windowed modules: rest sums), or null
  unsigned tie_xor;      // 0: equal bounds go oldest parent first; ~0u: newest first
  int nwmin;             // entries of wmin
  const unsigned long long *wmin;  // K = 1: the compaction's per-workgroup smallest keys, or null
  int lazy;              // K = 1 over a queue that keeps consumed (NaN) and ruled-out (lb > ub)
                         // entries between compactions: wmin always gives the box, stale
                         // entries are skipped, and the chunks to rescan are recorded
  int grow;              // batch slots with a larger tier above: a step that could exceed
                         // the queue (q_cap) or the parent arena stops with GPUGO_ST_GROW
                         // before anything is consumed, instead of overflowing
  long long q_cap;       // queue capacity (entries), read when grow is set
  unsigned W;            // wid^c children per split box, read when grow is set
  int grow_age;          // with grow: > 0 also stops a slot that has run this many steps
  BandView band;         // single solves with K > 1: the queue's priority band (st null: none)
};

// Selection key (lb, parent, pos), unique per box.  Parent records are
// numbered in creation order (collection keeps the order), so with
// tie_xor = ~0u equal bounds go to the newest parent first: depth-first
// among ties, which large-d problems with many equal bounds need.
__device__ __forceinline__ u128 qkey(const SelectArgs &a, long long i) {
  return ((u128)order_key(a.q_lb[i]) << 64) | ((u128)(a.q_par[i] ^ a.tie_xor) << 32) | (u128)a.q_pos[i];
}

// bin of a lower bound in [lmin, lmax]: monotone non-decreasing in lb
__device__ __forceinline__ unsigned value_bin(double lb, double lmin, double scale) {
  double t = (lb - lmin) * scale;
  int b = (int)t;
  return (unsigned)(b < 0 ? 0 : (b >= SEL_BINS ? SEL_BINS - 1 : b));
}

// Radix refinement on the 128-bit keys (12-bit digits) until the bin that
// holds the K-th smallest key has at most SEL_FINAL entries or fills the
// remaining slots exactly; then keys below that bin go to skey/sidx and the
// bin's keys to ckey/cidx.  Used when the value histogram cannot separate
// the keys (many equal lower bounds).
__device__ void radix_gather(const SelectArgs &a, long long N, int K, unsigned *hist, u128 *skey, unsigned *sidx,
                             u128 *ckey, unsigned *cidx, unsigned *scan_sh, u128 &s_pfx, int &s_pbits,
                             unsigned &s_krem, unsigned &s_region, unsigned &s_b, unsigned &s_cb, unsigned &s_ns,
                             unsigned &s_nc) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_pfx = 0;
    s_pbits = 0;
    s_krem = (unsigned)K;
  }
  __syncthreads();
  for (;;) {
    const int pbits = s_pbits;
    const int db = 128 - pbits < 12 ? 128 - pbits : 12;
    const u128 pfx = s_pfx;
    const unsigned krem = s_krem;
    const int sh = 128 - pbits - db;
    for (int b = tid; b < SEL_BINS; b += SEL_T) hist[b] = 0;
    __syncthreads();
    for (long long i = tid; i < N; i += SEL_T) {
      u128 k = qkey(a, i);
      bool m = (pbits == 0) || ((k >> (128 - pbits)) == (pfx >> (128 - pbits)));
      if (m) atomicAdd(&hist[(unsigned)(k >> sh) & ((1u << db) - 1u)], 1u);
    }
    __syncthreads();
    unsigned h[4], ts = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      h[q] = hist[4 * tid + q];
      ts += h[q];
    }
    unsigned tot;
    unsigned cum = block_excl_scan(ts, scan_sh, &tot);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (cum < krem && krem <= cum + h[q]) {
        s_b = 4 * tid + q;
        s_cb = cum;
        s_region = h[q];
      }
      cum += h[q];
    }
    __syncthreads();
    if (tid == 0) {
      s_krem = krem - s_cb;
      s_pfx = pfx | ((u128)s_b << sh);
      s_pbits = pbits + db;
    }
    __syncthreads();
    if (s_region == s_krem || s_region <= SEL_FINAL || s_pbits >= 128) break;
  }
  if (tid == 0) {
    s_ns = 0;
    s_nc = 0;
  }
  __syncthreads();
  const int shp = 128 - s_pbits;
  const u128 ptop = s_pfx >> shp;
  for (long long i = tid; i < N; i += SEL_T) {
    u128 k = qkey(a, i);
    u128 top = k >> shp;
    if (top < ptop) {
      unsigned slot = atomicAdd(&s_ns, 1u);
      if (slot < SEL_KMAX) {
        skey[slot] = k;
        sidx[slot] = (unsigned)i;
      }
    } else if (top == ptop) {
      unsigned slot = atomicAdd(&s_nc, 1u);
      if (slot < SEL_FINAL) {
        ckey[slot] = k;
        cidx[slot] = (unsigned)i;
      }
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// parent-record collection.  Records are appended, one per split box, and
// referenced by the queue's q_par; when the arena has no room for a step's
// new records, the records that no live entry references are dropped: the
// kept ones move down in index order and q_par is rewritten, so the
// (lb, parent, pos) order of the queue is unchanged and the solve goes on
// exactly as with a larger arena.  Works in chunks of GC_CHUNK records with
// a bitmap of referenced records and its word prefix counts in LDS (`bits`,
// `pre`: GC_WORDS words each).  Returns the number of records kept and
// remaps `par` (a parent index held in a register, or ~0u).
// ---------------------------------------------------------------------------
#define GC_WORDS SEL_BINS
#define GC_CHUNK (GC_WORDS * 32)
#define GC_U 4

__device__ __forceinline__ unsigned gc_rank(const unsigned *bits, const unsigned *pre, unsigned r) {
  return pre[r >> 5] + (unsigned)__builtin_popcount(bits[r >> 5] & ((1u << (r & 31)) - 1u));
}

__device__ long long collect_parents(const SelectArgs &a, long long N, long long used, unsigned *bits, unsigned *pre,
                                     unsigned *scan_sh, unsigned &par, double ub) {
  static_assert(GC_WORDS == 4 * SEL_T, "four bitmap words per thread");
  const int tid = threadIdx.x;
  const unsigned d = (unsigned)a.d;
  long long kept = 0;
  for (long long c0 = 0; c0 < used; c0 += GC_CHUNK) {
    const unsigned cn = (unsigned)(used - c0 < GC_CHUNK ? used - c0 : GC_CHUNK);
    for (int w = tid; w < GC_WORDS; w += SEL_T) bits[w] = 0u;
    __syncthreads();
    for (long long i0 = tid; i0 < N; i0 += (long long)SEL_T * SEL_U) {
      unsigned v[SEL_U];
#pragma unroll
      for (int u = 0; u < SEL_U; ++u) {
        const long long i = i0 + (long long)u * SEL_T;
        // a lazy queue's stale entries (NaN or lb > ub: never selected
        // again) reference nothing; their q_par is left as it is
        v[u] = (i < N && (!a.lazy || a.q_lb[i] <= ub)) ? a.q_par[i] : 0xffffffffu;
      }
#pragma unroll
      for (int u = 0; u < SEL_U; ++u) {
        const long long r = (long long)v[u] - c0;
        if (r >= 0 && r < (long long)cn) atomicOr(&bits[r >> 5], 1u << (r & 31));
      }
    }
    __syncthreads();
    unsigned h[4], ts = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      h[q] = (unsigned)__builtin_popcount(bits[4 * tid + q]);
      ts += h[q];
    }
    unsigned tot;
    unsigned cum = block_excl_scan(ts, scan_sh, &tot);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      pre[4 * tid + q] = cum;
      cum += h[q];
    }
    __syncthreads();
    // A kept record moves to an index at or below its own, and the flat
    // position r * pstride + j grows with the element number r * d + j, so
    // each range of elements is read whole before it is written and no
    // write reaches a later range.
    const unsigned ne = cn * d;
    for (unsigned e0 = 0; e0 < ne; e0 += SEL_T * GC_U) {
      double vl[GC_U], vh[GC_U], vw[GC_U];
      long long dst[GC_U];
#pragma unroll
      for (int u = 0; u < GC_U; ++u) {
        const unsigned e = e0 + (unsigned)u * SEL_T + tid;
        dst[u] = -1;
        if (e < ne) {
          const unsigned r = e / d, j = e - r * d;
          if ((bits[r >> 5] >> (r & 31)) & 1u) {
            const long long np = kept + gc_rank(bits, pre, r), op = c0 + r;
            if (np != op) {
              const long long s = op * a.pstride + j;
              vl[u] = a.plo[s];
              vh[u] = a.phi[s];
              vw[u] = a.pw[s];
              dst[u] = np * a.pstride + j;
            }
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < GC_U; ++u)
        if (dst[u] >= 0) {
          a.plo[dst[u]] = vl[u];
          a.phi[dst[u]] = vh[u];
          a.pw[dst[u]] = vw[u];
        }
    }
    for (unsigned r0 = 0; r0 < cn; r0 += SEL_T * GC_U) {
      int vn[GC_U];
      long long dst[GC_U];
#pragma unroll
      for (int u = 0; u < GC_U; ++u) {
        const unsigned r = r0 + (unsigned)u * SEL_T + tid;
        dst[u] = -1;
        if (r < cn && ((bits[r >> 5] >> (r & 31)) & 1u)) {
          const long long np = kept + gc_rank(bits, pre, r), op = c0 + r;
          if (np != op) {
            vn[u] = a.pwin[op];
            dst[u] = np;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < GC_U; ++u)
        if (dst[u] >= 0) a.pwin[dst[u]] = vn[u];
    }
    // per-record payload (ext_n doubles each), in the same element order
    const unsigned nx = cn * (unsigned)a.ext_n;
    for (unsigned e0 = 0; e0 < nx; e0 += SEL_T * GC_U) {
      double vx[GC_U];
      long long dst[GC_U];
#pragma unroll
      for (int u = 0; u < GC_U; ++u) {
        const unsigned e = e0 + (unsigned)u * SEL_T + tid;
        dst[u] = -1;
        if (e < nx) {
          const unsigned r = e / (unsigned)a.ext_n, k = e - r * (unsigned)a.ext_n;
          if ((bits[r >> 5] >> (r & 31)) & 1u) {
            const long long np = kept + gc_rank(bits, pre, r), op = c0 + r;
            if (np != op) {
              vx[u] = a.pext[op * a.ext_n + k];
              dst[u] = np * a.ext_n + k;
            }
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < GC_U; ++u)
        if (dst[u] >= 0) a.pext[dst[u]] = vx[u];
    }
    // every referenced index is below c0 + cn and maps at or below itself,
    // so rewritten entries are never taken up again by a later chunk
    for (long long i = tid; i < N; i += SEL_T) {
      if (a.lazy && !(a.q_lb[i] <= ub)) continue;
      const long long r = (long long)a.q_par[i] - c0;
      if (r >= 0 && r < (long long)cn) a.q_par[i] = (unsigned)(kept + gc_rank(bits, pre, (unsigned)r));
    }
    if (par != 0xffffffffu) {
      const long long r = (long long)par - c0;
      if (r >= 0 && r < (long long)cn) par = (unsigned)(kept + gc_rank(bits, pre, (unsigned)r));
    }
    kept += tot;
    __syncthreads();
  }
  return kept;
}


__device__ void select_body(const SelectArgs &a, const BandView *band = nullptr) {
  if (!band && a.band.st) band = &a.band;
  // the band's state, loaded with the problem's: the select is one
  // workgroup, and a load issued later would add a round trip to every step
  const long long band_n = band ? (long long)band->st->n : -1;
  const double band_g = band ? band->st->g : -__builtin_huge_val();
  __shared__ unsigned hist[SEL_BINS];
  __shared__ u128 skey[SEL_KMAX];
  __shared__ unsigned sidx[SEL_KMAX];
  __shared__ u128 ckey[SEL_FINAL];
  __shared__ unsigned cidx[SEL_FINAL];
  __shared__ unsigned scan_sh[SEL_T / 64 + 1];
  __shared__ u128 s_pfx;
  __shared__ int s_pbits;
  __shared__ unsigned s_krem, s_region, s_b, s_cb, s_ns, s_nc;
  __shared__ unsigned long long s_outmin;
  __shared__ double s_gnext;

  DevState *st = a.st;
  const int tid = threadIdx.x;
  if (st->status != GPUGO_ST_RUNNING) {
    if (tid == 0) st->compact_final = 0;
    return;
  }
  if (tid == 0) s_gnext = -__builtin_huge_val();
  if (st->fresh) {
    // a batch slot's first step splits the root record the host wrote
    __syncthreads();
    if (tid == 0) st->fresh = 0;
    return;
  }
  if (st->overflow || (a.max_iters > 0 && st->iterations >= a.max_iters)) {
    // stop; the enclosure is [smallest live lower bound, ub] (the queue is
    // compacted, so it holds no consumed entries), none after an overflow
    const long long N = st->compact_count;
    if (tid == 0) s_outmin = order_key(__builtin_huge_val());
    __syncthreads();
    unsigned long long mk_ = order_key(__builtin_huge_val());
    for (long long i = tid; i < N; i += SEL_T) {
      const unsigned long long k = order_key(a.q_lb[i]);
      mk_ = k < mk_ ? k : mk_;
    }
    atomicMin(&s_outmin, mk_);
    __syncthreads();
    if (tid == 0) {
      const bool ovf = st->overflow != 0;
      double lbq = key_value(s_outmin);
      if (st->out_min_lb < lbq) lbq = st->out_min_lb;
      if (ovf || (N == 0 && st->n_out == 0)) lbq = -__builtin_huge_val();
      st->lb_global = lbq;
      st->width = rnd<double, ROUND_EXACT>::add_up(key_value(st->ub_key), -lbq);
      st->status = ovf ? GPUGO_ST_OVERFLOW : GPUGO_ST_MAX_ITERS;
      st->q_count = N;
      st->compact_count = 0;
      st->compact_final = 1;
    }
    return;
  }
  const long long N = st->compact_count;
  const double ub = key_value(st->ub_key);
  if (N == 0) {
    if (tid == 0) {
      bool have_out = st->n_out > 0;
      st->status = have_out ? GPUGO_ST_X_CONVERGED : GPUGO_ST_EXHAUSTED;
      st->lb_global = have_out ? st->out_min_lb : -__builtin_huge_val();
      st->width = rnd<double, ROUND_EXACT>::add_up(ub, -st->lb_global);
      st->q_count = 0;
      st->compact_count = 0;
      st->compact_final = 1;
    }
    return;
  }
  const int K = (int)(N < (long long)a.k_select ? N : (long long)a.k_select);
  if (N <= K && !a.lazy) {
    for (long long i = tid; i < N; i += SEL_T) {
      skey[i] = qkey(a, i);
      sidx[i] = (unsigned)i;
    }
    __syncthreads();
  } else if (K == 1 && a.wmin) {
    // one box per step: the smallest of the compaction workgroups' smallest
    // keys (the same unique key order as the passes below), no queue pass
    if (tid < 64) {
      unsigned long long bh = ~0ull, bl = ~0ull, bo = 0ull;
      for (int q = tid; q < a.nwmin; q += 64) {
        const unsigned long long h = a.wmin[3 * q], l = a.wmin[3 * q + 1];
        if (h < bh || (h == bh && l < bl)) {
          bh = h;
          bl = l;
          bo = a.wmin[3 * q + 2];
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long th = __shfl_xor(bh, o, 64), tl = __shfl_xor(bl, o, 64), to = __shfl_xor(bo, o, 64);
        if (th < bh || (th == bh && tl < bl)) {
          bh = th;
          bl = tl;
          bo = to;
        }
      }
      if (tid == 0) {
        skey[0] = ((u128)bh << 64) | (u128)bl;
        sidx[0] = (unsigned)bo;
      }
    }
    __syncthreads();
    if (a.lazy) {
      // every entry is stale (ruled out since, or consumed): the live set is
      // empty, as in the N == 0 case above
      const unsigned long long kh = (unsigned long long)(skey[0] >> 64);
      if (kh == ~0ull || !(key_value(kh) <= ub)) {
        if (tid == 0) {
          bool have_out = st->n_out > 0;
          st->status = have_out ? GPUGO_ST_X_CONVERGED : GPUGO_ST_EXHAUSTED;
          st->lb_global = have_out ? st->out_min_lb : -__builtin_huge_val();
          st->width = rnd<double, ROUND_EXACT>::add_up(ub, -st->lb_global);
          st->q_count = 0;
          st->compact_count = 0;
          st->compact_final = 1;
        }
        return;
      }
    }
  } else {
    // Source of the selection: the slot's band when it is complete and holds
    // at least K entries, else the whole queue.
    const long long bn = band_n;
    const double bg = band_g;
    const bool use_band = bn >= (long long)K && bn <= band->cap && bg > -__builtin_huge_val();
    SelectArgs sv = a;
    long long Nsel = N;
    double lmax = key_value(st->lbmax_key);
    if (use_band) {
      sv.q_lb = const_cast<double *>(band->lb);
      sv.q_par = const_cast<unsigned *>(band->par);
      sv.q_pos = const_cast<unsigned *>(band->pos);
      Nsel = bn;
      if (bg < lmax) lmax = bg;  // every band entry is below g
    }
    if (tid == 0) s_gnext = bg;
    // Fast path: one histogram pass over the live range [lbmin, lbmax]
    // (tracked by the compaction).  The bin map is monotone in lb, so every
    // key in a bin below the boundary bin is smaller than every key in it.
    bool fast_done = false;
    const double lmin = key_value(st->lbmin_key);
    const double span = lmax - lmin;
    if (lmin > -__builtin_huge_val() && lmax < __builtin_huge_val() && span > 0) {
      const double scale = (double)SEL_BINS / span;
      for (int b = tid; b < SEL_BINS; b += SEL_T) hist[b] = 0;
      __syncthreads();
      // SEL_U loads in flight per thread: one workgroup's pass over the
      // queue is latency-bound otherwise
      for (long long i0 = tid; i0 < Nsel; i0 += (long long)SEL_T * SEL_U) {
        double v[SEL_U];
#pragma unroll
        for (int u = 0; u < SEL_U; ++u) {
          const long long i = i0 + (long long)u * SEL_T;
          v[u] = i < Nsel ? sv.q_lb[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < SEL_U; ++u)
          if (i0 + (long long)u * SEL_T < Nsel) atomicAdd(&hist[value_bin(v[u], lmin, scale)], 1u);
      }
      __syncthreads();
      unsigned h[4], ts = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        h[q] = hist[4 * tid + q];
        ts += h[q];
      }
      unsigned tot;
      unsigned cum = block_excl_scan(ts, scan_sh, &tot);
      const unsigned gt = (unsigned)BAND_TARGET > (unsigned)K ? (unsigned)BAND_TARGET : (unsigned)K;
      // a new threshold only for a queue larger than BAND_MIN (> gt); a
      // smaller one keeps no band: its passes are short, and copying all
      // of it to the band would cost the compaction more than they do
      const bool set_g = band && !use_band && tot > (unsigned)BAND_MIN;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (cum < (unsigned)K && (unsigned)K <= cum + h[q]) {
          s_b = 4 * tid + q;
          s_cb = cum;
          s_region = h[q];
        }
        if (set_g && cum < gt && gt <= cum + h[q])
          s_gnext = (4 * tid + q + 1 == SEL_BINS) ? __builtin_huge_val()
                                                   : lmin + (double)(4 * tid + q + 1) * (span / (double)SEL_BINS);
        cum += h[q];
      }
      if (band && !use_band && tid == 0 && !set_g) s_gnext = -__builtin_huge_val();
      __syncthreads();
      if (s_region <= SEL_FINAL || s_region == (unsigned)K - s_cb) {
        if (tid == 0) {
          s_ns = 0;
          s_nc = 0;
          s_krem = (unsigned)K - s_cb;
        }
        __syncthreads();
        const unsigned bb = s_b;
        for (long long i0 = tid; i0 < Nsel; i0 += (long long)SEL_T * SEL_U) {
          double v[SEL_U];
#pragma unroll
          for (int u = 0; u < SEL_U; ++u) {
            const long long i = i0 + (long long)u * SEL_T;
            v[u] = i < Nsel ? sv.q_lb[i] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < SEL_U; ++u) {
            const long long i = i0 + (long long)u * SEL_T;
            if (i >= Nsel) continue;
            const unsigned bin = value_bin(v[u], lmin, scale);
            if (bin < bb) {
              unsigned slot = atomicAdd(&s_ns, 1u);
              if (slot < SEL_KMAX) {
                skey[slot] = qkey(sv, i);
                sidx[slot] = (unsigned)i;
              }
            } else if (bin == bb) {
              unsigned slot = atomicAdd(&s_nc, 1u);
              if (slot < SEL_FINAL) {
                ckey[slot] = qkey(sv, i);
                cidx[slot] = (unsigned)i;
              }
            }
          }
        }
        __syncthreads();
        fast_done = true;
      }
    }
    if (!fast_done) radix_gather(sv, Nsel, K, hist, skey, sidx, ckey, cidx, scan_sh, s_pfx, s_pbits, s_krem, s_region,
                                 s_b, s_cb, s_ns, s_nc);
    // the s_krem smallest candidates fill the remaining slots (rank by
    // counting: keys are unique)
    const unsigned ns = s_ns, nc = s_nc, krem = s_krem;
    if (tid < (int)nc) {
      const u128 mine = ckey[tid];
      unsigned r = 0;
      for (unsigned j = 0; j < nc; ++j) r += ckey[j] < mine;
      if (r < krem) {
        skey[ns + r] = mine;
        sidx[ns + r] = cidx[tid];
      }
    }
    __syncthreads();
    if (use_band) {
      // band positions -> queue indices
      if (tid < K) sidx[tid] = band->idx[sidx[tid]];
      __syncthreads();
    }
  }
  // deterministic order of the selection: rank by counting into ckey/cidx
  {
    u128 mine = 0;
    unsigned myidx = 0, r = 0;
    if (tid < K) {
      mine = skey[tid];
      myidx = sidx[tid];
      for (int j = 0; j < K; ++j) r += skey[j] < mine;
    }
    __syncthreads();
    if (tid < K) {
      ckey[r] = mine;
      cidx[r] = myidx;
    }
    __syncthreads();
    if (tid < K) {
      skey[tid] = ckey[tid];
      sidx[tid] = cidx[tid];
    }
    __syncthreads();
  }
  // stop test: certified f-enclosure
  const double lb0 = a.q_lb[sidx[0]];
  const double lbg = lb0 < st->out_min_lb ? lb0 : st->out_min_lb;
  if (a.tol_f > 0) {
    double w = rnd<double, ROUND_EXACT>::add_up(ub, -lbg);
    if (w <= a.tol_f) {
      __syncthreads();
      if (tid == 0) {
        st->lb_global = lbg;
        st->width = w;
        st->status = GPUGO_ST_CERTIFIED;
        st->q_count = N;
        st->compact_count = 0;
        st->compact_final = 1;
      }
      return;
    }
  }

  // materialise: x-width output rule, then new parent records.  The
  // selected boxes' coordinates are handled one (box, coordinate) pair per
  // thread, so the cost does not grow with d per thread.  Per-box data
  // lives in the ranking buffers, which are free from here on.
  static_assert(sizeof(ckey) >= SEL_KMAX * (8 + 4 + 4) && SEL_FINAL >= SEL_KMAX, "per-box scratch fits in ckey");
  unsigned long long *bw = reinterpret_cast<unsigned long long *>(ckey);  // [K] max child width (bits)
  unsigned *bpar = reinterpret_cast<unsigned *>(bw + SEL_KMAX);            // [K] parent record
  unsigned *bpos = bpar + SEL_KMAX;                                         // [K] child index
  unsigned *bdst = cidx;                                                    // [K] destination
  int is_out = 0;
  unsigned idx = 0, par = 0, pos = 0;
  double elb = 0;
  const int d = a.d;
  const long long kd = (long long)K * d;
  if (tid < K) {
    idx = sidx[tid];
    par = a.q_par[idx];
    pos = a.q_pos[idx];
    elb = a.q_lb[idx];
    bpar[tid] = par;
    bpos[tid] = pos;
    bw[tid] = 0ull;
  }
  __syncthreads();
  if (a.tol_x > 0) {
    // max coordinate width of each selected child (widths are >= 0, so
    // their bit patterns order like the values)
    for (long long e = tid; e < kd; e += SEL_T) {
      const int b = (int)(e / d), j = (int)(e - (long long)b * d);
      const long long rec = (long long)bpar[b] * a.pstride;
      double lo, hi;
      child_coord(a.plo + rec, a.phi + rec, a.pw + rec, a.pwin[bpar[b]], d, a.wid, a.c, bpos[b], j, lo, hi);
      const double w = hi - lo;
      unsigned long long wb;
      __builtin_memcpy(&wb, &w, 8);
      atomicMax(&bw[b], wb);
    }
    __syncthreads();
    if (tid < K) {
      double m;
      const unsigned long long wb = bw[tid];
      __builtin_memcpy(&m, &wb, 8);
      is_out = m < a.tol_x;
    }
  }
  unsigned n_split, n_outs;
  const unsigned split_rank = block_excl_scan((tid < K && !is_out) ? 1u : 0u, scan_sh, &n_split);
  const unsigned out_rank = block_excl_scan((tid < K && is_out) ? 1u : 0u, scan_sh, &n_outs);
  if (a.grow && (N + (long long)n_split * (long long)a.W > a.q_cap ||
                 (a.grow_age > 0 && st->iterations >= (long long)a.grow_age))) {
    // this step's children might not fit the queue (or the slot has run its
    // steps in this tier): stop before anything is consumed, so the slot can
    // resume in a larger tier
    if (tid == 0) {
      st->status = GPUGO_ST_GROW;
      st->q_count = N;
      st->compact_final = 0;
    }
    return;
  }
  long long pbase = st->parents_used, room = a.p_cap;
  const long long obase = st->n_out;
  const bool collected = pbase + (long long)n_split > room;
  if (collected) {
    // arena full: drop the records no queue entry references (hist and skey
    // are free from here on); overflow if the kept ones and this step's
    // leave less than an eighth of the arena, so collections stay rare
    unsigned gpar = tid < K ? par : 0xffffffffu;
    pbase = collect_parents(a, N, pbase, hist, reinterpret_cast<unsigned *>(skey), scan_sh, gpar, ub);
    if (tid < K) par = gpar;
    room = a.p_cap - (a.p_cap >> 3);
  }
  if (pbase + (long long)n_split > room) {
    if (tid == 0) {
      if (a.grow) {
        // the records kept by the collection (renumbered, order kept) are
        // the slot's state; nothing was consumed
        st->status = GPUGO_ST_GROW;
        st->parents_used = pbase;
        st->q_count = N;
        st->compact_final = 0;
      } else {
        st->overflow |= 2;
        st->status = GPUGO_ST_OVERFLOW;
        st->q_count = N;
        st->compact_count = 0;
        st->compact_final = 1;
      }
    }
    return;
  }
  if (tid == 0) s_outmin = order_key(__builtin_huge_val());
  unsigned *bwin = hist;  // [K] window start of each box's parent (hist is free)
  if (tid < K) {
    bpar[tid] = par;  // remapped if the arena was collected
    bwin[tid] = (unsigned)a.pwin[par];
    bdst[tid] = is_out ? (0x80000000u | out_rank) : split_rank;
  }
  __syncthreads();
  // new parent records (split boxes) and output boxes (x-width rule): their
  // sources are kept records below pbase, their destinations at or above it
  for (long long e = tid; e < kd; e += SEL_T) {
    const int b = (int)(e / d), j = (int)(e - (long long)b * d);
    const long long rec = (long long)bpar[b] * a.pstride;
    double lo, hi;
    child_coord(a.plo + rec, a.phi + rec, a.pw + rec, (int)bwin[b], d, a.wid, a.c, bpos[b], j, lo, hi);
    const unsigned dst = bdst[b];
    if (!(dst & 0x80000000u)) {
      const long long np = (pbase + dst) * a.pstride + j;
      a.plo[np] = lo;
      a.phi[np] = hi;
      a.pw[np] = (hi - lo) / (double)a.wid;
    } else {
      const long long o = obase + (dst & 0x7fffffffu);
      if (o < a.out_cap) {
        a.out_lo[o * d + j] = lo;
        a.out_hi[o * d + j] = hi;
      }
    }
  }
  if (tid < K) {
    if (!is_out) {
      int nw = (int)bwin[tid] + a.c;
      while (nw >= d) nw -= d;
      a.pwin[pbase + split_rank] = nw;
    } else {
      const long long o = obase + out_rank;
      if (o < a.out_cap) a.out_lb[o] = elb;
      else atomicOr(&st->overflow, 4);
      atomicMin(&s_outmin, order_key(elb));
    }
    a.q_lb[idx] = __builtin_nan("");
  }
  __syncthreads();
  if (tid == 0) {
    double om = key_value(s_outmin);
    if (om < st->out_min_lb) st->out_min_lb = om;
    st->overflow &= ~4;  // output-arena overflow keeps the bound (out_min_lb) rigorous
    st->lb_global = lbg;
    st->first_parent = pbase;
    st->parents_used = pbase + n_split;
    st->n_split = (int)n_split;
    st->n_out = obase + n_outs;
    st->ub_snap = ub;
    st->q_count = N;
    st->compact_count = 0;
    st->lbmin_key = ~0ull;
    st->lbmax_key = 0ull;
    if (a.lazy) {
      // this step's appends start at N; a collection renumbered parent
      // records, which are part of the keys, so every chunk minimum is stale
      st->dirty_lo = collected ? 0 : (N < st->dirty_lo ? N : st->dirty_lo);
      st->consumed = idx;
    }
    if (band) {
      // this step's compaction rebuilds the band below the (new) threshold
      band->st->n = 0;
      band->st->g = K == 1 && a.wmin ? -__builtin_huge_val() : s_gnext;
    }
  }
}


// ---------------------------------------------------------------------------
// compact: keep entries with lb <= ub (consumed entries are NaN)
// ---------------------------------------------------------------------------
struct CompactArgs {
  DevState *st;
  const double *lb_in;
  const unsigned *par_in, *pos_in;
  double *lb_out;
  unsigned *par_out, *pos_out;
  long long cap;
  int *poll;             // or null: receives the problem's status (host-mapped pinned memory)
  unsigned tie_xor;      // key order of the next select (SelectArgs::tie_xor)
  unsigned long long *wmin;  // or null: [3 * nblk] smallest kept key (hi, lo) and its output index per workgroup
  unsigned long long *cmin;  // or null: chunk minima of a lazy queue, [3 * chunks] key (hi, lo) and index
  int chunk_log2;            // entries per chunk: 1 << chunk_log2
  int pad_c;
  BandView band;             // st null, or the band this pass rebuilds (kept entries with lb < band.st->g)
};

// blk / nblk: this workgroup's index and the number of workgroups sharing
// the pass (any block size up to 1024 threads)
// Entries per thread in one pass of a compaction workgroup: its loads are in
// flight together, and the pass costs one scan, one global atomic and two
// barriers for CMP_U * blockDim entries.
#define CMP_U 8

__device__ __forceinline__ void compact_body(const CompactArgs &a, int blk, int nblk) {
  // kept entries per (slab u, wavefront), then their exclusive prefix in
  // (u, wavefront) order: the output order of the pass
  __shared__ unsigned wcnt[CMP_U * 16];
  __shared__ unsigned long long s_base;
  DevState *st = a.st;
  // the host polls the status here instead of copying the state back
  if (a.poll && blk == 0 && threadIdx.x == 0)
    __hip_atomic_store(a.poll, st->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  // after a stop, exactly one more pass runs so the live set lands in the
  // buffer that st->compactions designates
  if (st->status != GPUGO_ST_RUNNING && !st->compact_final) return;
  __shared__ unsigned long long s_kmin, s_kmax;
  long long n = st->q_count;
  if (n > a.cap) n = a.cap;
  const double ub = key_value(st->ub_key);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // range of the kept lower bounds (order keys; feeds the next select)
  unsigned long long kmin = ~0ull, kmax = 0ull;
  if (threadIdx.x == 0) {
    s_kmin = ~0ull;
    s_kmax = 0ull;
  }
  const int nw = (int)(blockDim.x >> 6), bsz = (int)blockDim.x;
  const int nv = CMP_U * nw;  // <= 128 counters
  // smallest kept key (order_key(lb), (parent ^ tie_xor) << 32 | pos) and its
  // output index, for a one-box select that needs no pass over the queue
  unsigned long long mh = ~0ull, ml = ~0ull, mo = 0ull;
  // kept entries below the band threshold also go to the band
  BandState *bst = a.band.st;
  const double bg = bst ? bst->g : -__builtin_huge_val();
  const long long span = (long long)bsz * CMP_U;
  for (long long base = (long long)blk * span; base < n; base += (long long)nblk * span) {
    double lb[CMP_U];
#pragma unroll
    for (int u = 0; u < CMP_U; ++u) {
      const long long i = base + (long long)u * bsz + threadIdx.x;
      lb[u] = i < n ? a.lb_in[i] : 0.0;
    }
    unsigned kept = 0u, bbits = 0u, pre[CMP_U];
#pragma unroll
    for (int u = 0; u < CMP_U; ++u) {
      const long long i = base + (long long)u * bsz + threadIdx.x;
      const bool keep = i < n && lb[u] <= ub;
      if (keep) {
        const unsigned long long k = order_key(lb[u]);
        kmin = k < kmin ? k : kmin;
        kmax = k > kmax ? k : kmax;
        kept |= 1u << u;
        if (lb[u] < bg) bbits |= 1u << u;
      }
      const unsigned long long mask = __ballot(keep);
      pre[u] = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
      if (lane == 0) wcnt[u * nw + w] = (unsigned)__popcll(mask);
    }
    // band slots of this pass: one atomic per wavefront; the lane's entries
    // follow the lanes below it, in slab order
    unsigned long long bbase = 0;
    if (__ballot(bbits != 0u)) {
      const unsigned nb = (unsigned)__popc(bbits);
      const unsigned incl = wave_incl_scan(nb);
      const unsigned tot = __shfl(incl, 63, 64);
      unsigned long long b0 = 0;
      if (lane == 0) b0 = atomicAdd(&bst->n, (unsigned long long)tot);
      bbase = (unsigned long long)__shfl((long long)b0, 0, 64) + (incl - nb);
    }
    __syncthreads();
    if (w == 0) {
      // exclusive prefix of the nv counters (two per lane), one atomic
      const unsigned v0 = lane < nv ? wcnt[lane] : 0u, v1 = lane + 64 < nv ? wcnt[lane + 64] : 0u;
      const unsigned s0 = wave_incl_scan(v0), s1 = wave_incl_scan(v1);
      const unsigned t0 = __shfl(s0, 63, 64), t1 = __shfl(s1, 63, 64);
      if (lane < nv) wcnt[lane] = s0 - v0;
      if (lane + 64 < nv) wcnt[lane + 64] = t0 + s1 - v1;
      if (lane == 0) {
        const unsigned tot = t0 + t1;
        s_base = tot ? atomicAdd((unsigned long long *)&st->compact_count, (unsigned long long)tot) : 0ull;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CMP_U; ++u) {
      if (kept & (1u << u)) {
        const long long i = base + (long long)u * bsz + threadIdx.x;
        const long long o = (long long)s_base + wcnt[u * nw + w] + pre[u];
        const unsigned par = a.par_in[i], pos = a.pos_in[i];
        a.lb_out[o] = lb[u];
        a.par_out[o] = par;
        a.pos_out[o] = pos;
        if ((bbits >> u) & 1u) {
          const unsigned long long bo = bbase + (unsigned)__popc(bbits & ((1u << u) - 1u));
          if (bo < (unsigned long long)a.band.cap) {
            const_cast<double *>(a.band.lb)[bo] = lb[u];
            const_cast<unsigned *>(a.band.par)[bo] = par;
            const_cast<unsigned *>(a.band.pos)[bo] = pos;
            const_cast<unsigned *>(a.band.idx)[bo] = (unsigned)o;
          }
        }
        if (a.wmin) {
          const unsigned long long kh = order_key(lb[u]), kl = ((unsigned long long)(par ^ a.tie_xor) << 32) | pos;
          if (kh < mh || (kh == mh && kl < ml)) {
            mh = kh;
            ml = kl;
            mo = (unsigned long long)o;
          }
        }
      }
    }
    __syncthreads();
  }
  if (a.wmin) {
    __shared__ unsigned long long s_mk[16][3];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long th = __shfl_xor(mh, o, 64), tl = __shfl_xor(ml, o, 64), to = __shfl_xor(mo, o, 64);
      if (th < mh || (th == mh && tl < ml)) {
        mh = th;
        ml = tl;
        mo = to;
      }
    }
    if (lane == 0) {
      s_mk[w][0] = mh;
      s_mk[w][1] = ml;
      s_mk[w][2] = mo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < nw; ++q)
        if (s_mk[q][0] < mh || (s_mk[q][0] == mh && s_mk[q][1] < ml)) {
          mh = s_mk[q][0];
          ml = s_mk[q][1];
          mo = s_mk[q][2];
        }
      a.wmin[3 * blk] = mh;
      a.wmin[3 * blk + 1] = ml;
      a.wmin[3 * blk + 2] = mo;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long t0 = __shfl_xor(kmin, o, 64), t1 = __shfl_xor(kmax, o, 64);
    kmin = t0 < kmin ? t0 : kmin;
    kmax = t1 > kmax ? t1 : kmax;
  }
  if (lane == 0 && kmin <= kmax) {
    atomicMin(&s_kmin, kmin);
    atomicMax(&s_kmax, kmax);
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_kmin <= s_kmax) {
    atomicMin(&st->lbmin_key, s_kmin);
    atomicMax(&st->lbmax_key, s_kmax);
  }
  if (blk == 0 && threadIdx.x == 0) {
    atomicMax((unsigned long long *)&st->queue_peak, (unsigned long long)n);
    st->compactions += 1;
    st->dirty_lo = 0;  // entries moved: chunk minima (if kept) are stale
  }
}

// Smaller of two keys (hi, lo) with their indices.
__device__ __forceinline__ void key_min3(unsigned long long &h, unsigned long long &l, unsigned long long &o,
                                         unsigned long long th, unsigned long long tl, unsigned long long to) {
  if (th < h || (th == h && tl < l)) {
    h = th;
    l = tl;
    o = to;
  }
}

// Workgroup minimum of per-thread keys (any block size up to 1024); the
// result is valid in thread 0.
__device__ __forceinline__ void block_min3(unsigned long long &h, unsigned long long &l, unsigned long long &o) {
  __shared__ unsigned long long s_k[16][3];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
#pragma unroll
  for (int s = 32; s > 0; s >>= 1)
    key_min3(h, l, o, __shfl_xor(h, s, 64), __shfl_xor(l, s, 64), __shfl_xor(o, s, 64));
  __syncthreads();
  if (lane == 0) {
    s_k[w][0] = h;
    s_k[w][1] = l;
    s_k[w][2] = o;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 1; q < nw; ++q) key_min3(h, l, o, s_k[q][0], s_k[q][1], s_k[q][2]);
}

// ---------------------------------------------------------------------------
// lazy step tail (one box per step): instead of compacting, refresh the
// minima of the queue chunks that changed since the last select -- the
// chunks from st->dirty_lo up (this step's appends) and the chunk of the
// consumed entry -- keeping entries with lb <= ub, then reduce each
// workgroup's chunks (chunk ch belongs to workgroup ch mod nblk) into
// wmin[blk] for the select.  Consumed and ruled-out entries stay in place
// until the next compaction; they are never the minimum of a live queue
// (their keys are NaN or above ub, every live key is below).  After a stop
// the pass is the stop step's compaction, as compact_body.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lazy_chunks_body(const CompactArgs &a, int blk, int nblk) {
  __shared__ unsigned long long s_best[3];
  DevState *st = a.st;
  if (st->status != GPUGO_ST_RUNNING) {
    compact_body(a, blk, nblk);
    return;
  }
  if (a.poll && blk == 0 && threadIdx.x == 0)
    __hip_atomic_store(a.poll, st->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  long long n = st->q_count;
  if (n > a.cap) n = a.cap;
  const double ub = key_value(st->ub_key);
  const int sh = a.chunk_log2;
  const long long nch = (n + (1ll << sh) - 1) >> sh;
  const long long dlo = st->dirty_lo, cons = st->consumed;
  const long long cd = (dlo < n ? dlo : n) >> sh;       // first chunk with appends
  const long long cc = (cons >= 0 && cons < n) ? cons >> sh : -1;
  if (threadIdx.x == 0) {
    s_best[0] = ~0ull;
    s_best[1] = ~0ull;
    s_best[2] = 0ull;
  }
  for (long long ch = blk; ch < nch; ch += nblk) {
    if (ch < cd && ch != cc) continue;
    const long long e0 = ch << sh, e1 = (e0 + (1ll << sh) < n) ? e0 + (1ll << sh) : n;
    unsigned long long h = ~0ull, l = ~0ull, o = 0ull;
    for (long long i = e0 + threadIdx.x; i < e1; i += blockDim.x) {
      const double lb = a.lb_in[i];
      if (!(lb <= ub)) continue;
      const unsigned long long kh = order_key(lb);
      if (kh > h) continue;
      const unsigned long long kl = ((unsigned long long)(a.par_in[i] ^ a.tie_xor) << 32) | a.pos_in[i];
      key_min3(h, l, o, kh, kl, (unsigned long long)i);
    }
    block_min3(h, l, o);
    if (threadIdx.x == 0) {
      a.cmin[3 * ch] = h;
      a.cmin[3 * ch + 1] = l;
      a.cmin[3 * ch + 2] = o;
      key_min3(s_best[0], s_best[1], s_best[2], h, l, o);
    }
  }
  // the workgroup's other chunks are unchanged: their minima from earlier passes
  unsigned long long h = ~0ull, l = ~0ull, o = 0ull;
  for (long long ch = blk + (long long)threadIdx.x * nblk; ch < nch; ch += (long long)blockDim.x * nblk) {
    if (ch >= cd || ch == cc) continue;
    key_min3(h, l, o, a.cmin[3 * ch], a.cmin[3 * ch + 1], a.cmin[3 * ch + 2]);
  }
  block_min3(h, l, o);
  if (threadIdx.x == 0) {
    key_min3(h, l, o, s_best[0], s_best[1], s_best[2]);
    a.wmin[3 * blk] = h;
    a.wmin[3 * blk + 1] = l;
    a.wmin[3 * blk + 2] = o;
    if (blk == 0) {
      st->compact_count = n;
      atomicMax((unsigned long long *)&st->queue_peak, (unsigned long long)n);
    }
  }
}



// Barrier over all workgroups of a cooperative launch (every workgroup is
// resident).  bar[0] counts arrivals, bar[1] is the generation.  The wait is
// bounded: after ~2^24 sleeps it gives up and returns false (the caller
// stops and reports the error), so a mistake cannot hang the device.
// Thread 0 fences before arriving (release) and after the wait (acquire):
// the per-CU L1 and the per-XCD L2 are not coherent across the device, and
// a device-scope fence writes back / invalidates them, which is expensive
// on MI355X (8 XCDs), so one fence per workgroup and barrier, not per wave.
__device__ __forceinline__ bool grid_barrier(unsigned *bar, unsigned nblocks) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned g = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    if (atomicAdd(&bar[0], 1u) == nblocks - 1u) {
      __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned spin = 0;
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        __builtin_amdgcn_s_sleep(2);
        if (++spin > (1u << 24)) {
          ok = 0;
          break;
        }
      }
    }
    __threadfence();
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Arguments of the persistent step loop (a cooperative launch that runs
// select -> split+bound -> compact with grid barriers until the problem
// stops or max_steps steps ran).  The queue pointers of `ba`, `sa` are
// replaced per step by buffer (step & 1) of q_lb/q_par/q_pos.
struct PersistArgs {
  BoundArgs ba;
  SelectArgs sa;
  double *q_lb[2];
  unsigned *q_par[2], *q_pos[2];
  unsigned *bar;          // grid barrier state, zero at launch
  int *err;               // set to 1 if a barrier wait gave up
  long long step0;        // steps run before this launch (0: the first splits the root)
  long long max_steps;
  long long cap;          // queue capacity
  int nblocks, pad;
};

// Root of a solve: parent record 0 is the initial box, one box to split.
struct RootArgs {
  DevState *st;
  double *plo, *phi, *pw;
  int *pwin;
  const double *lo, *hi;
  double init_ub;
  int d, wid;
  BandState *band;  // or null: the problem's priority band, emptied with the state
};

__device__ __forceinline__ void root_body(const RootArgs &a) {
  if (a.band && threadIdx.x == 0) {
    a.band->g = -__builtin_huge_val();
    a.band->n = 0;
  }
  for (int j = threadIdx.x; j < a.d; j += blockDim.x) {
    a.plo[j] = a.lo[j];
    a.phi[j] = a.hi[j];
    a.pw[j] = (a.hi[j] - a.lo[j]) / (double)a.wid;
  }
  if (threadIdx.x == 0) {
    a.pwin[0] = 0;
    DevState s;
    __builtin_memset(&s, 0, sizeof s);
    s.ub_key = order_key(a.init_ub);
    s.ub_snap = a.init_ub;
    s.lb_global = -__builtin_huge_val();
    s.out_min_lb = __builtin_huge_val();
    s.width = __builtin_huge_val();
    s.parents_used = 1;
    s.first_parent = 0;
    s.n_split = 1;
    s.status = GPUGO_ST_RUNNING;
    s.lbmin_key = ~0ull;
    s.lbmax_key = 0ull;
    *a.st = s;
  }
}

}  // namespace gpugo

// queues up to this many entries get the stop's final compaction from the
// workgroup that stopped the problem; larger ones in a launch of their own
#ifndef GPUGO_FINAL_INLINE
#define GPUGO_FINAL_INLINE 65536
#endif
// bit of the polled status word: the stop's final compaction has not run
#define GPUGO_POLL_FINAL 0x100
// bit of the polled status word: the problem has stopped and the pinned copy
// of its final state (pstate) was written before the word
#define GPUGO_POLL_STATE 0x200

namespace gpugo {
// Compaction of this step fused with the select of the next one: every
// workgroup compacts its share; the last to finish (ticket in st->cblocks,
// each workgroup fences its writes before taking one, the last fences
// before reading) selects from the compacted queue.  Saves one launch and
// its gap per step.  `poll` (or null) receives the status after the select;
// once the problem has stopped and its final compaction has run, `pstate`
// (host-mapped pinned memory, or null) first receives the whole final state,
// so the host needs no copy back.
__device__ __forceinline__ void compact_select_body(const CompactArgs &ca, const SelectArgs &sa, int *poll,
                                                    DevState *pstate) {
  __shared__ int s_last;
  if (ca.cmin) lazy_chunks_body(ca, blockIdx.x, gridDim.x);
  else compact_body(ca, blockIdx.x, gridDim.x);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const int t = atomicAdd(&ca.st->cblocks, 1);
    s_last = t == (int)gridDim.x - 1;
    if (s_last) {
      ca.st->cblocks = 0;
      if (ca.cmin) {
        // every workgroup has read these; all chunk minima are current now
        ca.st->dirty_lo = 0x7fffffffffffffffll;
        ca.st->consumed = -1;
      }
      __threadfence();
    }
  }
  __syncthreads();
  if (!s_last) return;
  select_body(sa);
  __syncthreads();
  // A stop with a small queue: this workgroup runs the stop's final
  // compaction itself (from the buffer the select read into the other one),
  // so the host needs no launch after the stop.
  __shared__ int s_fin;
  if (threadIdx.x == 0) {
    const DevState *st = ca.st;
    s_fin = st->status != GPUGO_ST_RUNNING && st->compact_final && st->q_count <= GPUGO_FINAL_INLINE;
  }
  __syncthreads();
  if (s_fin) {
    CompactArgs fa = ca;
    if (!ca.cmin) {
      // a full pass wrote the out buffer and the select read it
      fa.lb_in = ca.lb_out;
      fa.par_in = ca.par_out;
      fa.pos_in = ca.pos_out;
      // (both buffers are the problem's own writable queue arrays)
      fa.lb_out = const_cast<double *>(ca.lb_in);
      fa.par_out = const_cast<unsigned *>(ca.par_in);
      fa.pos_out = const_cast<unsigned *>(ca.pos_in);
    }
    fa.wmin = nullptr;
    fa.cmin = nullptr;
    fa.poll = nullptr;
    fa.band.st = nullptr;  // the problem has stopped: no select reads a band again
    compact_body(fa, 0, 1);
    __syncthreads();
    if (threadIdx.x == 0) ca.st->compact_final = 0;
    __syncthreads();
  }
  // the host reads the status, with GPUGO_POLL_FINAL if the stop's final
  // compaction is still to run
  if (poll && threadIdx.x < 64) {
    // wavefront 0: its lanes read the state words together
    const int status = __hip_atomic_load(&ca.st->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int cfinal = __hip_atomic_load(&ca.st->compact_final, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int word = status | (cfinal ? GPUGO_POLL_FINAL : 0);
    if (pstate && status != GPUGO_ST_RUNNING && !cfinal) {
      // the state no longer changes: every workgroup has finished its pass.
      // Lane i copies word i: device-scope loads (this CU's L1 may hold
      // words that other workgroups, or this one's atomics, changed since),
      // system-scope stores to the pinned copy.
      constexpr int NW = (int)(sizeof(DevState) / 8);
      static_assert(sizeof(DevState) % 8 == 0 && NW <= 64, "DevState is copied in 8-byte words by one wavefront");
      const int i = (int)threadIdx.x;
      if (i < NW) {
        unsigned long long *src = reinterpret_cast<unsigned long long *>(ca.st);
        unsigned long long *dst = reinterpret_cast<unsigned long long *>(pstate);
        __hip_atomic_store(dst + i, __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      word |= GPUGO_POLL_STATE;
    }
    // release (the wavefront's stores above land first), then the word the
    // host waits for
    if (threadIdx.x == 0) __hip_atomic_store(poll, word, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
}  // namespace gpugo

#ifdef GPUGO_MODULE_STEP
extern "C" __global__ __launch_bounds__(SEL_T) void gpugo_compact_select(gpugo::CompactArgs ca, gpugo::SelectArgs sa,
                                                                         int *poll, gpugo::DevState *pstate) {
  gpugo::compact_select_body(ca, sa, poll, pstate);
}
extern "C" __global__ __launch_bounds__(SEL_T) void gpugo_select(gpugo::SelectArgs a) { gpugo::select_body(a); }
extern "C" __global__ __launch_bounds__(256) void gpugo_compact(gpugo::CompactArgs a) {
  gpugo::compact_body(a, blockIdx.x, gridDim.x);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_root(gpugo::RootArgs a) { gpugo::root_body(a); }
#endif
