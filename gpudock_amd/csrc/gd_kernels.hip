/*
 * gd_kernels.hip -- CDNA4 (gfx950) kernels of the REMC docking core.
 *
 *   gd_mc_kernel       fused Move + CalcEnergy + Accept for one cycle of MC
 *                      steps of every replica of every resident complex
 *                      (montecarlo_d.cu:59-106 with move_d.cu:73-134,
 *                      calcenergy_d.cu:22-374, accept_d.cu:12-52), or the
 *                      initial energies (MonteCarlo_Init_d, :12-55).
 *   gd_mc_split_kernel the same cycle with one replica split over 4, 8 or
 *                      16 wavefronts (launches with few replicas).
 *   gd_exchange1_kernel / gd_exchange_kernel temperature swaps, one
 *                      wavefront per chain, one lane per temperature pair
 *                      (exchangereplicas_d.cu:7-136), then the advance.
 *   gd_advance_kernel  advances each complex's device step counter after a
 *                      cycle and re-arms the work queue.
 *   gd_score_kernel    energies of caller-given poses (FP32).
 *   gd_rescore64_kernel FP64 energies of caller-given poses.
 *   gd_rescore64_best_kernel FP64 energies of the best poses of listed
 *                      replicas of many complexes (result read-back).
 *
 * Execution model: a workgroup takes a task -- up to WPG replicas of one
 * protein conformer of one complex -- either by its block index or, in the
 * persistent form, from a device work queue (an atomic counter over a task
 * list sorted by cost).  It stages that conformer's pocket points, the
 * pair-parameter rows of the ligand's atom types and (when it fits) the
 * type-grouped KDE points in LDS, restaging only when the next task needs
 * another conformer; each wavefront then owns one replica for the whole
 * cycle.  Lanes run over pocket points (over KDE points / ligand atoms in
 * the smaller loops); the ligand atom of the outer loop is broadcast from
 * the lane that holds it with v_readlane, so it sits in scalar registers.
 * The replica's uniform state (movematrices, energies) lives in a per-wave
 * LDS slot, not in vector registers; read-only uniform data (the task, pack
 * fields, each atom's table row, type and charge) is read through the
 * scalar unit, and the per-atom PSP rows and HBM-resident KDE points through
 * raw buffer resources with a fixed lane offset and a scalar chunk offset.  Per-term sums are per-lane FP32
 * accumulators combined with the near-first pairing tree the CPU replay
 * reproduces bit for bit.  The step index of each complex lives in device
 * memory, so the three launches of a cycle (MC, exchange, advance) replay
 * unchanged from a captured hipGraph.
 */
#include <hip/hip_runtime.h>
#include <mutex>
#include "gd_device.h"
#include "gd_launch.h"

/* Word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data
 * format, no swizzle): reads past num_records return 0. */
#define GD_BUFFER_DWORD3 0x00020000

typedef float gd_f4 __attribute__((ext_vector_type(4)));
typedef float gd_f2 __attribute__((ext_vector_type(2)));

template <bool B>
struct gd_bool {
  static constexpr bool value = B;
};

/* The wave's slot in LDS, addressed as LDS (not through flat accesses). */
typedef __attribute__((address_space(3))) float gd_lds_f32;

/* A read of read-only, wave-uniform data through the scalar unit. */
template <typename T>
__device__ __forceinline__ T gd_uload(const T *p)
{
  return *(const __attribute__((address_space(4))) T *) p;
}

/* One atom record through the scalar unit (a single 32-byte load). */
typedef int gd_i8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ gd_atom_rec gd_load_atom(const gd_atom_rec *p)
{
  return __builtin_bit_cast(gd_atom_rec, *(const __attribute__((address_space(4))) gd_i8 *) p);
}

__device__ __forceinline__ float gd_bcast(float v, int lane)
{
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <int CTRL>
__device__ __forceinline__ float gd_dpp(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF,
                                                    false));
}

/* Sum over the 64 lanes with the pairing tree of an xor butterfly taken
 * near-first (lane pairs j^1, then j^2, j^4, j^8, j^16, j^32): DPP
 * quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror (pairs the two quads
 * of a half-row), row_mirror (pairs the two halves of a row) leave every
 * lane of row i with the row sum r_i; row_bcast:15 adds r0 to row 1 and r2
 * to row 3, row_bcast:31 adds row 1's (r0 + r1) to row 3's (r2 + r3), read
 * from lane 63 (what the other rows receive is not read).  Every partial is the same float addition
 * as in that butterfly (addition commutes), so the result is bit-identical
 * to it; it is returned uniform. */
__device__ __forceinline__ float gd_wave_sum(float v)
{
  v += gd_dpp<0xB1>(v);
  v += gd_dpp<0x4E>(v);
  v += gd_dpp<0x141>(v);
  v += gd_dpp<0x140>(v);
  v += gd_dpp<0x142>(v);   /* rows 1 and 3 are the ones read */
  v += gd_dpp<0x143>(v);   /* row 3 is the one read */
  return gd_bcast(v, 63);
}

__device__ __forceinline__ double gd_wave_sum64(double v)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
    v += __shfl_xor(v, off, GD_WAVE);
  return v;
}

/* ------------------------------------------------------------------------ */
/* Workgroup staging                                                         */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ void gd_stage(const gd_pack &P, int prt, char *smem)
{
  float4 *s_tab = (float4 *) (smem + P.off_tab);
  const int pnp_pad = P.pnp_pad;
  const float4 *g_pos = P.pt_pos + (size_t) prt * pnp_pad;
  const float2 *g_eh = P.pt_eh + (size_t) prt * pnp_pad;
  /* chunk c of 64 points: {x, y} of its points, then {z, type}, then
   * {ele, hpp}, 512 B each */
  if (P.pocket_in_lds)
    for (int i = threadIdx.x; i < pnp_pad; i += blockDim.x) {
      const float4 q = g_pos[i];
      float2 *c = (float2 *) (smem + (i / GD_WAVE) * GD_LDS_CHUNK_BYTES) + (i & (GD_WAVE - 1));
      c[0] = make_float2(q.x, q.y);
      c[GD_WAVE] = make_float2(q.z, q.w);
      c[2 * GD_WAVE] = g_eh[i];
    }
  const float4 *g_tab = (const float4 *) P.tab;
  const int ntab4 = P.n_rows * GD_ROW_STRIDE * 2;
  for (int i = threadIdx.x; i < ntab4; i += blockDim.x)
    s_tab[i] = g_tab[i];
  if (P.kde_in_lds) {
    float4 *s_kde = (float4 *) (smem + P.off_kde);
    const int n_kde = P.pnk + 1;   /* with the zero-weight padding point */
    for (int i = threadIdx.x; i < n_kde; i += blockDim.x)
      s_kde[i] = P.kde[i];
  }
}

/* Split launches (one replica per workgroup) with split_extra: the
 * (atom, KDE point) pair list, the restraint targets and the PSP rows of
 * protein conformer prt, read every step by every item, staged once. */
__device__ __forceinline__ void gd_stage_split_extra(const gd_pack &P, int prt, char *smem)
{
  const int n_words = ((P.kde_n_lo + GD_WAVE - 1) / GD_WAVE +
                       (P.kde_n_hi + GD_WAVE - 1) / GD_WAVE) * GD_WAVE;
  uint32_t *s_pairs = (uint32_t *) (smem + P.off_pairs);
  for (int i = threadIdx.x; i < n_words; i += blockDim.x)
    s_pairs[i] = P.kde_pairs[i];
  float4 *s_mcs = (float4 *) (smem + P.off_mcs);
  const int n_mcs = P.pos * P.lna;
  for (int i = threadIdx.x; i < n_mcs; i += blockDim.x)
    s_mcs[i] = P.mcs[i];
  float *s_psp = (float *) (smem + P.off_psp);
  const float *g_psp = P.psp_pt + (size_t) prt * P.n_rows * P.pnp_pad;
  const int n_psp = P.n_rows * P.pnp_pad;
  for (int i = threadIdx.x; i < n_psp; i += blockDim.x)
    s_psp[i] = g_psp[i];
}

/* ------------------------------------------------------------------------ */
/* Energy of one pose (calcenergy_d.cu:22-374), in summation blocks          */
/* (gd_block_size in gd_canon.h): atom blocks of the pair loop, blocks of    */
/* KDE rounds, one item per position restraint.  Lane j holds ligand atoms   */
/* j and j + 64 in (ax0..az0) and (ax1..az1).  prt = the protein conformer   */
/* staged in smem.                                                           */
/* ------------------------------------------------------------------------ */
/* Lane partials of the directly summed terms over one atom block. */
struct gd_blk {
  float vdw, pmf, psp, hdb, ele;
};

/* Pair loop + hydrophobic restraint (calcenergy_d.cu:56-149) over ligand
 * atoms [l0, l1): lane partials from +0 into blk, the block's restraint sum
 * (uniform) into *ehpc_b, MCC confusion counts (uniform) added to tp/fp/fn.
 * XL (split launches, whose complexes all stage the extras): pocket points,
 * KDE points and PSP rows are all in LDS; else the PSP rows come from HBM
 * and the pocket from LDS or HBM as staged. */
template <bool XL>
__device__ __forceinline__ void gd_atoms_block(
    const gd_pack &P, const char *smem, int prt, float ax0, float ay0, float az0, float ax1,
    float ay1, float az1, int l0, int l1, gd_blk &blk, float *ehpc_b, int &tp, int &fp,
    int &fn)
{
  const gd_pair_param *__restrict__ s_tab = (const gd_pair_param *) (smem + P.off_tab);
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const int lna = P.lna;
  const int nchunk = P.pnp_pad / GD_WAVE;   /* padding points add exact zeros */
  const int n_rows = P.n_rows;
  const gd_pair_scalars sc = P.sc;
  /* the near-field coefficients and lj0 sit beside other scalars in one
   * fma (one scalar operand per instruction): held in vector registers, the
   * loop needs no copy per pair */
  gd_pair_scalars sc_v = sc;
  asm volatile("" : "+v"(sc_v.ea1), "+v"(sc_v.eb1), "+v"(sc_v.lj0));

  gd_pair_acc acc;
  acc.vdw = 0.0f; acc.ele = 0.0f; acc.pmf = 0.0f; acc.psp = 0.0f; acc.hdb = 0.0f;
  float eele = 0.0f;
  float ehpc = 0.0f;
  /* MCC confusion counts (wave-uniform): contacts now vs reference, from one
   * ballot per chunk against the reference mask, read through scalar loads */
  typedef const __attribute__((address_space(4))) unsigned long long *mask_ptr;
  const mask_ptr mref =
      P.mcc_ref ? (mask_ptr) (P.mcc_ref + (size_t) prt * lna * nchunk) : nullptr;
  typedef const __attribute__((address_space(1))) float *gfloat_ptr;
  const gfloat_ptr psp_prt = (gfloat_ptr) (P.psp_pt + (size_t) prt * n_rows * P.pnp_pad);
  const int kp = gd_uload(P.chunk_bounds + 2 * prt + 1);
  const unsigned ulane = (unsigned) lane;
  const int pocket_lds = P.pocket_in_lds;
  const __amdgpu_buffer_rsrc_t pos_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void *) (P.pt_pos + (size_t) prt * P.pnp_pad), 0, P.pnp_pad * 16, GD_BUFFER_DWORD3);
  const __amdgpu_buffer_rsrc_t eh_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void *) (P.pt_eh + (size_t) prt * P.pnp_pad), 0, P.pnp_pad * 8, GD_BUFFER_DWORD3);
  const gd_lds_f32 *psp_lds = (const gd_lds_f32 *) (smem + P.off_psp);

  /* the atom loops, for one way of reading the pocket: the choice is made
   * once per call, not per atom */
  auto atoms = [&](auto read) {
    /* atoms l < 64 sit in lane l of (ax0, ay0, az0), atoms l >= 64 in lane
     * l - 64 of (ax1, ay1, az1): one pass per register set, so an atom's
     * coordinates are three readlanes */
#pragma unroll 1
    for (int g = l0 / GD_WAVE; g * GD_WAVE < l1; ++g) {
      const float gx = g ? ax1 : ax0, gy = g ? ay1 : ay0, gz = g ? az1 : az0;
      const int la = l0 > g * GD_WAVE ? l0 : g * GD_WAVE;
      const int lb = l1 < (g + 1) * GD_WAVE ? l1 : (g + 1) * GD_WAVE;
      for (int l = la; l < lb; ++l) {
        const int src = l & (GD_WAVE - 1);
        const float lx = gd_bcast(gx, src);
        const float ly = gd_bcast(gy, src);
        const float lz = gd_bcast(gz, src);
        const gd_atom_rec ar = gd_load_atom(P.atom + l);
        const int row = ar.row;
        const float lq = ar.q;
        const gd_pair_param *trow = s_tab + row * GD_ROW_STRIDE;
        /* the PSP row as a buffer: lane offset fixed, chunk offset scalar */
        const __amdgpu_buffer_rsrc_t psp_rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void *) (psp_prt + (size_t) row * P.pnp_pad), 0, P.pnp_pad * 4, GD_BUFFER_DWORD3);
        float hacc = 0.0f;
        acc.ele = 0.0f;
        /* chunks [0, kh) hold the hydrophobic-class points (the canonical order
         * puts them first) and [0, kp) the class-2 points, kp >= kh; beyond kp
         * every hydrophobic weight and every PSP value is 0, whose terms would
         * add exact zeros, so that loop leaves the terms (and the PSP read) out */
        auto chunks = [&](auto read, auto hpc_tag, auto hdb_tag, auto psp_tag, int k0, int k1) {
          for (int k = k0; k < k1; ++k) {
            float4 q;
            float2 eh;
            read(k, q, eh);
            const uint32_t bits = __float_as_uint(q.w);
            /* the PSP read (chunk base uniform, lane offset fixed) is consumed
             * at the end of the pair; points outside class 2 hold PSP 0 */
            float pv = 0.0f;
            if (decltype(psp_tag)::value) {
              if (XL)
                pv = psp_lds[row * P.pnp_pad + k * GD_WAVE + lane];
              else
                pv = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                    psp_rsrc, ulane * 4u, k * GD_WAVE * 4, 0));
            }
            const float d2 = gd_dist2(lx - q.x, ly - q.y, lz - q.z);
            if (decltype(hpc_tag)::value)
              hacc = gd_hpc_acc2(d2, eh.y, hacc);
            const gd_pair_param pp =
                *(const gd_pair_param *) ((const char *) trow + bits);
            const int near_any = __ballot(d2 < sc_v.dcut2) != 0;
            const int contact = gd_pair_terms(d2, eh.x, &pp, &sc_v, decltype(psp_tag)::value,
                                              pv, decltype(hdb_tag)::value, near_any, &acc);
            if (mref) {
              const unsigned long long now = __ballot(contact);
              const unsigned long long ref = mref[(size_t) l * nchunk + k];
              tp += __popcll(now & ref);
              fp += __popcll(now & ~ref);
              fn += __popcll(~now & ref);
            }
          }
        };
        /* rows of ligand types without any H-bond pair leave that test out;
         * the chunks [kh, kp) (none in what the loaders build) run with the
         * hydrophobic term, whose points there hold hpp = 0 and add exact
         * zeros, so an atom needs two loops, not three */
        auto rows = [&](auto read) {
          const gd_bool<true> on;
          const gd_bool<false> off;
          if (ar.hdb) {
            chunks(read, on, on, on, 0, kp);
            chunks(read, off, on, off, kp, nchunk);
          } else {
            chunks(read, on, off, on, 0, kp);
            chunks(read, off, off, off, kp, nchunk);
          }
        };
        rows(read);
        eele = gd_fma(lq, acc.ele, eele);
        const float hs = 0.5f * gd_wave_sum(hacc);   /* the pairs added twice the kernel */
        const float z = (hs - ar.hpl0) * ar.inv_hpl1;
        ehpc += 0.5f * z * z - ar.hpl2;
      }
    }
  };
  if (XL || pocket_lds) {
    atoms([&](int k, float4 &q, float2 &eh) {
      const float2 *__restrict__ c = (const float2 *) (smem + k * GD_LDS_CHUNK_BYTES) + lane;
      const float2 xy = c[0], zw = c[GD_WAVE];
      q = make_float4(xy.x, xy.y, zw.x, zw.y);
      eh = c[2 * GD_WAVE];
    });
  } else {
    /* whole-protein pockets: the conformer's points as raw buffers, lane
     * offset fixed, chunk offset scalar */
    atoms([&](int k, float4 &q, float2 &eh) {
      const gd_f4 v = __builtin_bit_cast(gd_f4, __builtin_amdgcn_raw_buffer_load_b128(
                                                    pos_rsrc, ulane * 16u, k * GD_WAVE * 16, 0));
      const gd_f2 w = __builtin_bit_cast(gd_f2, __builtin_amdgcn_raw_buffer_load_b64(
                                                    eh_rsrc, ulane * 8u, k * GD_WAVE * 8, 0));
      q = make_float4(v.x, v.y, v.z, v.w);
      eh = make_float2(w.x, w.y);
    });
  }
  blk.vdw = acc.vdw;
  blk.pmf = acc.pmf;
  blk.psp = acc.psp;
  blk.hdb = acc.hdb;
  blk.ele = eele;
  *ehpc_b = ehpc;
}

/* KDE rounds (calcenergy_d.cu:175-213): the pose's (atom, KDE point) pairs as
 * one list, 64 to a round -- atoms < 64 first, then atoms >= 64, each group
 * padded to whole rounds -- so that no lane idles at the end of an atom's
 * short segment.  Lane j of round r takes pair 64 r + j: its atom's
 * coordinates come from the lane that holds the atom, the point carries the
 * pair's weight 1 / (segment length) in .w; the padding pairs name a point
 * of weight 0 after the segments, so they add exact zeros (the sums are
 * never -0) and every lane keeps its sum.  Returns this lane's sum over
 * rounds [r0, r1) from +0. */
template <bool XL>
__device__ __forceinline__ float gd_kde_rounds(const gd_pack &P, const char *smem, float ax0,
                                               float ay0, float az0, float ax1, float ay1,
                                               float az1, int r0, int r1)
{
  const unsigned ulane = threadIdx.x & (GD_WAVE - 1);
  const float kde2 = P.kde2;
  const int n_lo = P.kde_n_lo, n_hi = P.kde_n_hi;
  const int rounds_lo = (n_lo + GD_WAVE - 1) / GD_WAVE;
  const int n_all = (rounds_lo + (n_hi + GD_WAVE - 1) / GD_WAVE) * GD_WAVE;
  const __amdgpu_buffer_rsrc_t pairs_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void *) P.kde_pairs, 0, n_all * 4, GD_BUFFER_DWORD3);
  float lane_sum = 0.0f;
  auto rounds = [&](auto read) {
    /* rounds [ra, rb) of one group, whose atoms lanes hold in (gx, gy, gz);
     * padding entries name lane 0 and the zero-weight point */
    auto entry = [&](int r) -> uint32_t {
      return XL ? ((const __attribute__((address_space(3))) uint32_t *) (smem + P.off_pairs))
                      [r * GD_WAVE + ulane]
                : __builtin_amdgcn_raw_buffer_load_b32(pairs_rsrc, ulane * 4u,
                                                       r * GD_WAVE * 4, 0);
    };
    auto group = [&](int ra, int rb, float gx, float gy, float gz) {
      if (ra >= rb)
        return;
      /* each round's entries are loaded one round ahead (the last round
       * loads its own again), so the point read does not wait for them */
      uint32_t e_next = entry(ra);
      for (int r = ra; r < rb; ++r) {
        const uint32_t e = e_next;
        e_next = entry(r + 1 < rb ? r + 1 : r);
        const int src = (int) (e & 0xffu);
        const float lx = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(gx)));
        const float ly = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(gy)));
        const float lz = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(gz)));
        const gd_f4 q = read(e >> 12);   /* byte offset of the point */
        const float d2 = gd_dist2(lx - q.x, ly - q.y, lz - q.z);
        float es;
        const float ep = gd_exp_hot_split(kde2 * d2, &es);
        lane_sum = gd_fma(ep * es, q.w, lane_sum);   /* ep * 2^n is exact */
      }
    };
    group(r0, r1 < rounds_lo ? r1 : rounds_lo, ax0, ay0, az0);
    group(r0 > rounds_lo ? r0 : rounds_lo, r1, ax1, ay1, az1);
  };
  if (XL || P.kde_in_lds) {
    const __attribute__((address_space(3))) char *pts =
        (const __attribute__((address_space(3))) char *) (smem + P.off_kde);
    rounds([&](uint32_t off) {
      return *(const __attribute__((address_space(3))) gd_f4 *) (pts + off);
    });
  } else {
    /* from HBM as a raw buffer */
    const __amdgpu_buffer_rsrc_t kde_rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void *) P.kde, 0, (P.pnk + 1) * 16, GD_BUFFER_DWORD3);
    rounds([&](uint32_t off) {
      return __builtin_bit_cast(gd_f4, __builtin_amdgcn_raw_buffer_load_b128(kde_rsrc, off, 0,
                                                                             0));
    });
  }
  return lane_sum;
}

/* KDE rounds of the pose: (rounds_lo + rounds_hi). */
__device__ __forceinline__ int gd_kde_round_count(const gd_pack &P)
{
  return (P.kde_n_lo + GD_WAVE - 1) / GD_WAVE + (P.kde_n_hi + GD_WAVE - 1) / GD_WAVE;
}

/* Position restraint m (calcenergy_d.cu:240-258): tcc_m sqrt(mean squared
 * distance of the atoms it names), 0 when it names none (uniform). */
template <bool XL>
__device__ __forceinline__ float gd_lhm_one(const gd_pack &P, const char *smem, float ax0,
                                            float ay0, float az0, float ax1, float ay1,
                                            float az1, int m)
{
  typedef const __attribute__((address_space(3))) gd_f4 *lds_f4;
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const int lna = P.lna;
  float s = 0.0f, c = 0.0f;
  const float4 *tm = P.mcs + (size_t) m * lna;
  const lds_f4 tl = (lds_f4) (smem + P.off_mcs) + m * lna;
  auto target = [&](int a) -> float4 {
    if (XL) {
      const gd_f4 v = tl[a];
      return make_float4(v.x, v.y, v.z, v.w);
    }
    return tm[a];
  };
  if (lane < lna) {
    const float4 t = target(lane);
    if (t.w != 0.0f) {
      s += gd_dist2(ax0 - t.x, ay0 - t.y, az0 - t.z);
      c += 1.0f;
    }
  }
  if (lane + GD_WAVE < lna) {
    const float4 t = target(lane + GD_WAVE);
    if (t.w != 0.0f) {
      s += gd_dist2(ax1 - t.x, ay1 - t.y, az1 - t.z);
      c += 1.0f;
    }
  }
  s = gd_wave_sum(s);
  c = gd_wave_sum(c);
  return c > 0.0f ? gd_uload(P.mcs_tcc + m) * gd_sqrt(s / c) : 0.0f;
}

/* From the summed terms to the energies: KDE normalisation, the restraint
 * log, the distance term, averaging, normalisation and total
 * (calcenergy_d.cu:276-371); out[] (slot order) and *cmcc uniform.  vdw ..
 * kde are the wave sums of the lane totals. */
/* distance of the placed centre to the pocket centre (calcenergy_d.cu:276-283) */
__device__ __forceinline__ float gd_edst(const gd_pack &P, const float cen[3])
{
  return gd_sqrt(gd_dist2(cen[0] - P.pocket_center[0], cen[1] - P.pocket_center[1],
                          cen[2] - P.pocket_center[2]));
}

__device__ __forceinline__ void gd_energy_final_d(const gd_pack &P, float vdw, float ele,
                                                  float pmf, float hdb, float psp, float ehpc,
                                                  float kde, float elhm_sum, int tp, int fp,
                                                  int fn, float edst, float out[GD_E_N],
                                                  float *cmcc)
{
  const int pos = P.pos;
  const float ekde = kde / P.kde3;
  const float elhm = (pos > 0) ? gd_log(elhm_sum / (float) pos) : 0.0f;
  float sums[GD_F_NTERM];
  sums[GD_F_VDW] = vdw;
  sums[GD_F_ELE] = ele * P.sc.inv_el1;   /* the pair loop sums el1 g */
  sums[GD_F_PMF] = pmf;
  sums[GD_F_HPC] = ehpc;
  sums[GD_F_HDB] = hdb;
  sums[GD_F_DST] = edst;
  sums[GD_F_PSP] = psp;
  sums[GD_F_KDE] = ekde;
  sums[GD_F_LHM] = elhm;
  gd_finalize(sums, (float) P.lna, P.a_para, P.b_para, P.w, P.term_order, P.combine, out);
  *cmcc = 0.0f;
  if (P.mcc_ref) {
    const int tn = P.lna * P.pnp - tp - fp - fn;   /* padding points are never contacts */
    *cmcc = gd_mcc((float) tp, (float) fn, (float) fp, (float) tn);
  }
}

__device__ __forceinline__ void gd_energy_final(const gd_pack &P, float vdw, float ele,
                                                float pmf, float hdb, float psp, float ehpc,
                                                float kde, float elhm_sum, int tp, int fp,
                                                int fn, const float cen[3], float out[GD_E_N],
                                                float *cmcc)
{
  gd_energy_final_d(P, vdw, ele, pmf, hdb, psp, ehpc, kde, elhm_sum, tp, fp, fn,
                    gd_edst(P, cen), out, cmcc);
}

/* The same from the lane totals: their wave sums first. */
__device__ __forceinline__ void gd_energy_tail(const gd_pack &P, const gd_blk &tot,
                                               float ehpc, float kde_lane, float elhm_sum,
                                               int tp, int fp, int fn, const float cen[3],
                                               float out[GD_E_N], float *cmcc)
{
  gd_energy_final(P, gd_wave_sum(tot.vdw), gd_wave_sum(tot.ele), gd_wave_sum(tot.pmf),
                  gd_wave_sum(tot.hdb), gd_wave_sum(tot.psp), ehpc, gd_wave_sum(kde_lane),
                  elhm_sum, tp, fp, fn, cen, out, cmcc);
}

/* One wavefront evaluates every block in order.  On return out[] (slot
 * order) and *cmcc (MCC against the reference contacts, CalcMcc_d
 * calcmcc_d.cu:148-213, or 0) are identical in every lane. */
__device__ __forceinline__ void gd_energy_wave(
    const gd_pack &P, const char *smem, int prt, float ax0, float ay0, float az0, float ax1,
    float ay1, float az1, const float cen[3], float out[GD_E_N], float *cmcc)
{
  gd_blk tot;
  tot.vdw = 0.0f; tot.pmf = 0.0f; tot.psp = 0.0f; tot.hdb = 0.0f; tot.ele = 0.0f;
  float ehpc = 0.0f;
  int tp = 0, fp = 0, fn = 0;
  const int lna = P.lna;
  const int bs = gd_block_size(lna);
  for (int l0 = 0; l0 < lna; l0 += bs) {
    gd_blk blk;
    float eb;
    gd_atoms_block<false>(P, smem, prt, ax0, ay0, az0, ax1, ay1, az1, l0, min(l0 + bs, lna),
                          blk, &eb, tp, fp, fn);
    tot.vdw += blk.vdw;
    tot.pmf += blk.pmf;
    tot.psp += blk.psp;
    tot.hdb += blk.hdb;
    tot.ele += blk.ele;
    ehpc += eb;
  }
  const int n_rounds = gd_kde_round_count(P);
  const int kbs = gd_block_size(n_rounds > 0 ? n_rounds : 1);
  float kde_lane = 0.0f;
  for (int r0 = 0; r0 < n_rounds; r0 += kbs)
    kde_lane += gd_kde_rounds<false>(P, smem, ax0, ay0, az0, ax1, ay1, az1, r0,
                                     min(r0 + kbs, n_rounds));
  float elhm = 0.0f;
  for (int m = 0; m < P.pos; ++m)
    elhm += gd_lhm_one<false>(P, smem, ax0, ay0, az0, ax1, ay1, az1, m);
  gd_energy_tail(P, tot, ehpc, kde_lane, elhm, tp, fp, fn, cen, out, cmcc);
}

/* Move the conformer held by this lane (move_d.cu:86-132). */
__device__ __forceinline__ void gd_place(const gd_pack &P, const float m[6],
                                         float bx0, float by0, float bz0,
                                         float bx1, float by1, float bz1,
                                         float *ax0, float *ay0, float *az0,
                                         float *ax1, float *ay1, float *az1,
                                         float cen[3])
{
  float rot[9];
  gd_rotation(m, rot);
  const float c[3] = {P.center[0], P.center[1], P.center[2]};
  gd_transform(rot, m, c, bx0, by0, bz0, ax0, ay0, az0);
  gd_transform(rot, m, c, bx1, by1, bz1, ax1, ay1, az1);
  cen[0] = c[0] + m[0];
  cen[1] = c[1] + m[1];
  cen[2] = c[2] + m[2];
}

/* Centroid-relative conformer coordinates, two atoms per lane. */
__device__ __forceinline__ void gd_load_conformer(const gd_pack &P, int conf, float *bx0,
                                                  float *by0, float *bz0, float *bx1,
                                                  float *by1, float *bz1)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const float4 *cx = P.lig_xyz + (size_t) conf * P.lna;
  *bx0 = *by0 = *bz0 = *bx1 = *by1 = *bz1 = 0.0f;
  if (lane < P.lna) {
    const float4 v = cx[lane];
    *bx0 = v.x; *by0 = v.y; *bz0 = v.z;
  }
  if (lane + GD_WAVE < P.lna) {
    const float4 v = cx[lane + GD_WAVE];
    *bx1 = v.x; *by1 = v.y; *bz1 = v.z;
  }
}

/* Steps of this cycle for a complex: spe while steps are left (steps_run is
 * a whole number of cycles), else 0. */
__device__ __forceinline__ int gd_cycle_steps(const gd_pack &P, int s0)
{
  const int left = P.steps_run - s0;
  return left < P.spe ? left : P.spe;
}

/* ------------------------------------------------------------------------ */
/* One replica, one wavefront.                                               */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ void gd_replica_init(const gd_pack &P, const char *smem, int r)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const gd_state S = P.S;
  float bx0, by0, bz0, bx1, by1, bz1;
  gd_load_conformer(P, P.rep_lig[r], &bx0, &by0, &bz0, &bx1, &by1, &bz1);
  float m[6];
#pragma unroll
  for (int k = 0; k < 6; ++k)
    m[k] = S.mv[r * 8 + k];
  if (P.away) {
#pragma unroll
    for (int k = 0; k < 6; ++k)
      m[k] = P.move_scale[k] * 44.5f + m[k];
  }
  float ax0, ay0, az0, ax1, ay1, az1, cen[3], e[GD_E_N], cm;
  gd_place(P, m, bx0, by0, bz0, bx1, by1, bz1, &ax0, &ay0, &az0, &ax1, &ay1, &az1, cen);
  gd_energy_wave(P, smem, gd_uload(P.rep_prt + r), ax0, ay0, az0, ax1, ay1, az1, cen, e, &cm);
  if (lane == 0) {
    S.cmcc[r] = cm;
    S.cmcc_best[r] = cm;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      S.mv[r * 8 + k] = m[k];
      S.mv_best[r * 8 + k] = m[k];
    }
#pragma unroll
    for (int i = 0; i < GD_E_N; ++i) {
      S.e[r * 10 + i] = e[i];
      S.e_best[r * 10 + i] = e[i];
    }
    S.step_best[r] = -1;
  }
}

/* ------------------------------------------------------------------------ */
/* One replica on SPLIT wavefronts.  Its area in LDS (floats): lane partials */
/* of every atom block (5 terms x 64 lanes) and of every block of KDE rounds */
/* (64 lanes); per atom block the restraint sum and the MCC counts; per      */
/* position restraint its value; three pose buffers (placed atoms, their   */
/* movematrix and centre); the summed terms; an item counter and the index  */
/* of the pose being evaluated.  The wavefronts take items (atom blocks,     */
/* then KDE blocks, then  */
/* restraints) from the counter until none is left, so one that finishes    */
/* early takes more.  Each item writes its own slots and the sums read them  */
/* in block order -- the order one wavefront keeps alone -- so the energies  */
/* have the same bits whichever wavefront evaluated which item.              */
/* ------------------------------------------------------------------------ */
#define GD_SA_PART 0
#define GD_SA_KPART (GD_MAX_BLOCKS * 5 * GD_WAVE)
#define GD_SA_EHPC (GD_SA_KPART + GD_MAX_BLOCKS * GD_WAVE)
#define GD_SA_MCC (GD_SA_EHPC + GD_MAX_BLOCKS)
#define GD_SA_LHM (GD_SA_MCC + GD_MAX_BLOCKS * 3)
/* pose buffer i: atoms k * 64 + lane (x0 y0 z0 x1 y1 z1), movematrix at
 * GD_PB_M, centre at GD_PB_CEN */
#define GD_PB_M (6 * GD_WAVE)
#define GD_PB_CEN (GD_PB_M + 6)
#define GD_PB_FLOATS (6 * GD_WAVE + 16)
#define GD_SA_POSE (GD_SA_LHM + GD_SPLIT_MAX_POS)
#define GD_SA_CONF (GD_SA_POSE + 3 * GD_PB_FLOATS)  /* the conformer, by lane */
#define GD_SA_SUMS (GD_SA_CONF + 6 * GD_WAVE)       /* the summed terms */
#define GD_SA_SEL (GD_SA_SUMS + 14)               /* the pose buffer of this step (int) */
#define GD_SA_CTR (GD_SA_SUMS + 15)               /* the item counter (int) */
#define GD_SA_FLOATS (GD_SA_SUMS + 16)
static_assert(GD_SA_FLOATS * 4 == GD_SPLIT_AREA_BYTES, "split area size");

typedef __attribute__((address_space(3))) int gd_lds_i32;

/* Item counts of a complex, read once per cycle (not after every barrier). */
struct gd_split_sizes {
  int bs, nb;          /* atom block size and count */
  int n_rounds, kbs, nbk;   /* KDE rounds, rounds per block, blocks */
  int pos, n_items;
  int mcc;             /* 1: MCC counts are kept */
};

__device__ __forceinline__ gd_split_sizes gd_split_sizes_of(const gd_pack &P)
{
  gd_split_sizes z;
  z.bs = gd_block_size(P.lna);
  z.nb = (P.lna + z.bs - 1) / z.bs;
  z.n_rounds = gd_kde_round_count(P);
  z.kbs = gd_block_size(z.n_rounds > 0 ? z.n_rounds : 1);
  z.nbk = (z.n_rounds + z.kbs - 1) / z.kbs;
  z.pos = P.pos;
  z.n_items = z.nb + z.nbk + z.pos;
  z.mcc = P.mcc_ref != nullptr;
  return z;
}

template <bool XL>
__device__ __forceinline__ void gd_energy_items(
    const gd_pack &P, const char *smem, gd_lds_f32 *A, int prt, const gd_split_sizes &Z,
    float ax0, float ay0, float az0, float ax1, float ay1, float az1)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const int lna = P.lna;
  const int bs = Z.bs, nb = Z.nb, n_rounds = Z.n_rounds, kbs = Z.kbs, nbk = Z.nbk;
  const int n_items = Z.n_items;
  gd_lds_i32 *ctr = (gd_lds_i32 *) (A + GD_SA_CTR);
  for (;;) {
    /* every lane adds 1 (one LDS atomic of 64 for the wavefront) and the
     * first lane's old value / 64 is the item: no lane-divergent branch
     * feeds the loop */
    const int got = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
    const int i = __builtin_amdgcn_readfirstlane(got) >> 6;
    if (i >= n_items)
      break;
    if (i < nb) {
      gd_blk blk;
      float eb;
      int tp = 0, fp = 0, fn = 0;
      gd_atoms_block<XL>(P, smem, prt, ax0, ay0, az0, ax1, ay1, az1, i * bs,
                         min(i * bs + bs, lna), blk, &eb, tp, fp, fn);
      gd_lds_f32 *pb = A + GD_SA_PART + i * 5 * GD_WAVE + lane;
      pb[0 * GD_WAVE] = blk.vdw;
      pb[1 * GD_WAVE] = blk.pmf;
      pb[2 * GD_WAVE] = blk.psp;
      pb[3 * GD_WAVE] = blk.hdb;
      pb[4 * GD_WAVE] = blk.ele;
      if (lane == 0) {
        A[GD_SA_EHPC + i] = eb;
        A[GD_SA_MCC + 3 * i + 0] = __int_as_float(tp);
        A[GD_SA_MCC + 3 * i + 1] = __int_as_float(fp);
        A[GD_SA_MCC + 3 * i + 2] = __int_as_float(fn);
      }
    } else if (i < nb + nbk) {
      const int kb = i - nb;
      A[GD_SA_KPART + kb * GD_WAVE + lane] =
          gd_kde_rounds<XL>(P, smem, ax0, ay0, az0, ax1, ay1, az1, kb * kbs,
                            min(kb * kbs + kbs, n_rounds));
    } else {
      const int m = i - nb - nbk;
      const float v = gd_lhm_one<XL>(P, smem, ax0, ay0, az0, ax1, ay1, az1, m);
      if (lane == 0)
        A[GD_SA_LHM + m] = v;
    }
  }
}

/* t + src[0] + src[stride] + ... + src[(n - 1) stride], in that order, for
 * n <= GD_MAX_BLOCKS: the values are read eight at a time before they are
 * added, so the LDS loads overlap. */
__device__ __forceinline__ float gd_ordered_sum(const gd_lds_f32 *src, int stride, int n)
{
  float t = 0.0f;
#pragma unroll
  for (int b0 = 0; b0 < GD_MAX_BLOCKS; b0 += 8) {
    if (b0 >= n)
      break;
    float v[8];
#pragma unroll
    for (int b = 0; b < 8; ++b)
      v[b] = src[(b0 + b) * stride];
#pragma unroll
    for (int b = 0; b < 8; ++b)
      if (b0 + b < n)
        t += v[b];
  }
  return t;
}

/* After the items: the partials added in block order, spread over the
 * wavefronts -- job j on wavefront j % split: jobs 0-4 the vdw, pmf, psp,
 * hdb, ele lane totals over the atom blocks and their wave sums, job 5 the
 * KDE lane total and its wave sum, job 6 the restraint sum and the MCC
 * counts, job 7 the position-restraint sum -- into the sums area.  Every
 * job reads all its slots before adding (the area has GD_MAX_BLOCKS of
 * each; the adds use the first nb / nbk in order). */
__device__ __forceinline__ void gd_energy_sums(gd_lds_f32 *A, int role, int split,
                                               const gd_split_sizes &Z)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const int nb = Z.nb, nbk = Z.nbk;
  for (int j = role; j < 8; j += split) {
    if (j < 6) {
      const float t = gd_wave_sum(
          j < 5 ? gd_ordered_sum(A + GD_SA_PART + j * GD_WAVE + lane, 5 * GD_WAVE, nb)
                : gd_ordered_sum(A + GD_SA_KPART + lane, GD_WAVE, nbk));
      if (lane == 0)
        A[GD_SA_SUMS + j] = t;
    } else if (j == 6) {
      const float ehpc = gd_ordered_sum(A + GD_SA_EHPC, 1, nb);
      /* the MCC counts are integers: any order */
      int tp = 0, fp = 0, fn = 0;
      if (Z.mcc) {
        const bool in_b = lane < nb;
        tp = in_b ? __float_as_int(A[GD_SA_MCC + 3 * lane + 0]) : 0;
        fp = in_b ? __float_as_int(A[GD_SA_MCC + 3 * lane + 1]) : 0;
        fn = in_b ? __float_as_int(A[GD_SA_MCC + 3 * lane + 2]) : 0;
#pragma unroll
        for (int off = 1; off < GD_MAX_BLOCKS; off <<= 1) {
          tp += __shfl_xor(tp, off, GD_WAVE);
          fp += __shfl_xor(fp, off, GD_WAVE);
          fn += __shfl_xor(fn, off, GD_WAVE);
        }
      }
      if (lane == 0) {
        A[GD_SA_SUMS + 6] = ehpc;
        A[GD_SA_SUMS + 8] = __int_as_float(tp);
        A[GD_SA_SUMS + 9] = __int_as_float(fp);
        A[GD_SA_SUMS + 10] = __int_as_float(fn);
      }
    } else {
      /* position restraints in order, value m in lane m */
      const int pos = Z.pos;
      float elhm = 0.0f;
      if (pos > 0) {
        const float lv = lane < pos ? A[GD_SA_LHM + lane] : 0.0f;
        for (int m = 0; m < pos; ++m)
          elhm += gd_bcast(lv, m);
      }
      if (lane == 0)
        A[GD_SA_SUMS + 7] = elhm;
    }
  }
}

/* What the energies need of a complex, read once per cycle: lane k < 10
 * holds a_para[k] and b_para[k]; the weights are uniform. */
struct gd_final_consts {
  float a_l, b_l;
  float w[GD_F_NTERM];   /* uniform */
  float lna_f, kde3, inv_el1;
  int pos, term_order, combine;
};

__device__ __forceinline__ gd_final_consts gd_final_consts_of(const gd_pack &P)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  gd_final_consts F;
  F.a_l = lane < 10 ? P.a_para[lane] : 0.0f;
  F.b_l = lane < 10 ? P.b_para[lane] : 0.0f;
#pragma unroll
  for (int k = 0; k < GD_F_NTERM; ++k)
    F.w[k] = P.w[k];
  F.lna_f = (float) P.lna;
  F.kde3 = P.kde3;
  F.inv_el1 = P.sc.inv_el1;
  F.pos = P.pos;
  F.term_order = P.term_order;
  F.combine = P.combine;
  return F;
}

/* Wavefront 0, after the sums: the energies (edst: the distance term of the
 * placed centre).  The weighted-sum total (combine 0) with one term per
 * lane: lane L takes parameter L and the term it applies to -- file-order
 * term L (term_order 0) or the term of slot L (term_order 1) -- and forms
 * it with the operations of gd_energy_final_d / gd_finalize (one division
 * sequence serves every term); the total adds w x value over the lanes in
 * parameter order, as gd_finalize does.  Bitwise equal to
 * gd_energy_final_d, which the Bayesian total still uses. */
__device__ __forceinline__ void gd_energy_combine(const gd_pack &P, const gd_lds_f32 *A,
                                                  const gd_final_consts &F, float edst,
                                                  float out[GD_E_N], float *cmcc)
{
  const gd_lds_f32 *S = A + GD_SA_SUMS;
  if (F.combine != 0) {
    gd_energy_final_d(P, S[0], S[4], S[1], S[3], S[2], S[6], S[5], S[7],
                      __float_as_int(S[8]), __float_as_int(S[9]), __float_as_int(S[10]), edst,
                      out, cmcc);
    return;
  }
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const int L = lane < GD_F_NTERM ? lane : 0;
  /* file-order term of each slot (slot i holds file term s2f[i]), 4 bits each */
  const unsigned long long s2f = 0x587346210ull;
  /* the sums-area word of each file-order term (term 5 is edst) */
  const unsigned long long fword = 0x752036140ull;
  const int f = F.term_order ? (int) ((s2f >> (4 * L)) & 0xfu) : L;
  float x = S[(int) ((fword >> (4 * f)) & 0xfu)];
  if (f == GD_F_ELE)
    x = x * F.inv_el1;   /* the pair loop sums el1 g */
  if (f == GD_F_KDE)
    x = x / F.kde3;
  float y = x / F.lna_f;
  if (f == GD_F_HDB)
    y = y / 0x1.40d932p+1f * -1.0f;   /* sqrt(2 pi) */
  if (f == GD_F_DST)
    y = edst;
  if (f == GD_F_LHM)
    y = F.pos > 0 ? gd_log(x / (float) F.pos) : 0.0f;
  const float v = gd_fma(F.a_l, y, F.b_l);
  float total = 0.0f;
#pragma unroll
  for (int k = 0; k < GD_F_NTERM; ++k)
    total = gd_fma(F.w[k], gd_bcast(v, k), total);
#pragma unroll
  for (int i = 0; i < GD_F_NTERM; ++i) {
    /* the lane that formed slot i */
    const int src = F.term_order ? i : (int) ((s2f >> (4 * i)) & 0xfu);
    out[i] = gd_bcast(v, src);
  }
  out[GD_E_TOTAL] = total;
  *cmcc = 0.0f;
  if (P.mcc_ref) {
    const int tp = __float_as_int(S[8]), fp = __float_as_int(S[9]), fn = __float_as_int(S[10]);
    const int tn = P.lna * P.pnp - tp - fp - fn;
    *cmcc = gd_mcc((float) tp, (float) fn, (float) fp, (float) tn);
  }
}

/* The draws of step s of replica r (move_d.cu:73-84, accept_d.cu:20-27): the
 * movematrix displacement d[6] and the two acceptance draws (the second only
 * for the reversed test). */
__device__ __forceinline__ void gd_draws(const gd_pack &P, int r, int s, uint32_t seed,
                                         uint32_t cid, float d[6], float u_acc[2])
{
  const gd_u32x4 ra = gd_philox4x32_10((uint32_t) r, (uint32_t) s, GD_CTR_MOVE_A, 0u, seed, cid);
  const gd_u32x4 rb = gd_philox4x32_10((uint32_t) r, (uint32_t) s, GD_CTR_MOVE_B, 0u, seed, cid);
  const int fx = P.fixed_draws;
  d[0] = P.move_scale[0] * gd_perturbation(fx, ra.v[0]);
  d[1] = P.move_scale[1] * gd_perturbation(fx, ra.v[1]);
  d[2] = P.move_scale[2] * gd_perturbation(fx, ra.v[2]);
  d[3] = P.move_scale[3] * gd_perturbation(fx, ra.v[3]);
  d[4] = P.move_scale[4] * gd_perturbation(fx, rb.v[0]);
  d[5] = P.move_scale[5] * gd_perturbation(fx, rb.v[1]);
  u_acc[0] = gd_u01(rb.v[2]);
  u_acc[1] = gd_u01(rb.v[3]);
}

/* RecordLigand_d (util_d.cu:72-94): the state before step s, one word per
 * lane. */
__device__ __forceinline__ void gd_record(const gd_pack &P, const gd_lds_f32 *ws, int r,
                                          int itmp, int s)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  uint32_t *rec = P.rec + ((size_t) r * P.rec_per_window +
                           (s % P.rec_window) / P.rec_stride) * GD_REC_WORDS;
  float fv = 0.0f;
  if (lane >= 4 && lane < 14)
    fv = ws[GD_WS_EOLD + lane - 4];
  else if (lane == 14)
    fv = ws[GD_WS_CMCC];
  else if (lane >= 15 && lane < 21)
    fv = ws[GD_WS_MOLD + lane - 15];
  uint32_t v = __float_as_uint(fv);
  if (lane == 0) v = (uint32_t) r;
  if (lane == 1) v = (uint32_t) P.rep_prt[r];
  if (lane == 2) v = (uint32_t) itmp;
  if (lane == 3) v = (uint32_t) P.rep_lig[r];
  if (lane == 21) v = (uint32_t) s;
  if (lane < GD_REC_WORDS)
    rec[lane] = v;
}

/* Metropolis (accept_d.cu:12-52) on wavefront-uniform energies: the state
 * in the slot, the best pose in HBM; returns 1 if the move was taken. */
__device__ __forceinline__ int gd_accept(const gd_pack &P, gd_lds_f32 *ws, int r, int s,
                                         const float u_acc[2], float mb,
                                         const float e_new[GD_E_N], float cm_new,
                                         float *e_old_total, float *best_total)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const gd_state S = P.S;
  if (!gd_metropolis(e_new[GD_E_TOTAL], *e_old_total, mb, u_acc[0], u_acc[1], P.fixed_draws,
                     P.reverse))
    return 0;
  *e_old_total = e_new[GD_E_TOTAL];
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < GD_E_N; ++i)
      ws[GD_WS_EOLD + i] = e_new[i];
#pragma unroll
    for (int k = 0; k < 6; ++k)
      ws[GD_WS_MOLD + k] = ws[GD_WS_MNEW + k];
    ws[GD_WS_CMCC] = cm_new;
  }
  if (e_new[GD_E_TOTAL] < *best_total) {
    *best_total = e_new[GD_E_TOTAL];
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 6; ++k)
        S.mv_best[r * 8 + k] = ws[GD_WS_MNEW + k];
#pragma unroll
      for (int i = 0; i < GD_E_N; ++i)
        S.e_best[r * 10 + i] = e_new[i];
      S.cmcc_best[r] = cm_new;
      S.step_best[r] = s;
    }
  }
  return 1;
}

/* The slot from the replica's state in HBM (wavefront 0 of the replica). */
__device__ __forceinline__ void gd_slot_load(const gd_pack &P, gd_lds_f32 *ws, int r)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const gd_state S = P.S;
  if (lane < 6)
    ws[GD_WS_MOLD + lane] = S.mv[r * 8 + lane];
  if (lane < GD_E_N)
    ws[GD_WS_EOLD + lane] = S.e[r * 10 + lane];
  if (lane == 0)
    ws[GD_WS_CMCC] = S.cmcc[r];
}

/* The replica's state in HBM from the slot, and the cycle's accept count. */
__device__ __forceinline__ void gd_slot_store(const gd_pack &P, const gd_lds_f32 *ws, int r,
                                              int acs)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const gd_state S = P.S;
  if (lane < 6)
    S.mv[r * 8 + lane] = ws[GD_WS_MOLD + lane];
  if (lane < GD_E_N)
    S.e[r * 10 + lane] = ws[GD_WS_EOLD + lane];
  if (lane == 0) {
    S.acs_mc[r] += acs;
    S.cmcc[r] = ws[GD_WS_CMCC];
  }
}

/* n_steps MC steps from step s0 of replica r on one wavefront.  The
 * movematrices and energies live in the replica's LDS slot
 * ws[GD_WSLOT_FLOATS]; only totals stay in registers. */
__device__ __forceinline__ void gd_replica_cycle(const gd_pack &P, const char *smem,
                                                 gd_lds_f32 *ws, int r, int prt, int s0,
                                                 int n_steps)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const gd_state S = P.S;
  const int lconf = gd_uload(P.rep_lig + r);
  gd_slot_load(P, ws, r);
  __builtin_amdgcn_wave_barrier();
  float e_old_total = S.e[r * 10 + GD_E_TOTAL];
  float best_total = S.e_best[r * 10 + GD_E_TOTAL];
  const int itmp = S.idx_tmp[r];
  const float mb = P.minus_beta[itmp];
  const uint32_t seed = P.seed, cid = P.cid;
  int acs = 0;

  for (int s = s0; s < s0 + n_steps; ++s) {
    if (P.rec != nullptr && s % P.rec_stride == 0)
      gd_record(P, ws, r, itmp, s);
    float d[6], u_acc[2];
    gd_draws(P, r, s, seed, cid, d, u_acc);
    float m_new[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
      m_new[k] = d[k] + ws[GD_WS_MOLD + k];

    /* the conformer is re-read every step (an L1 hit) rather than held in
     * registers across the energy evaluation */
    float bx0, by0, bz0, bx1, by1, bz1;
    gd_load_conformer(P, lconf, &bx0, &by0, &bz0, &bx1, &by1, &bz1);
    float ax0, ay0, az0, ax1, ay1, az1, cen[3], e_new[GD_E_N], cm_new;
    gd_place(P, m_new, bx0, by0, bz0, bx1, by1, bz1, &ax0, &ay0, &az0, &ax1, &ay1,
             &az1, cen);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 6; ++k)
        ws[GD_WS_MNEW + k] = m_new[k];
    }
    gd_energy_wave(P, smem, prt, ax0, ay0, az0, ax1, ay1, az1, cen, e_new, &cm_new);
    acs += gd_accept(P, ws, r, s, u_acc, mb, e_new, cm_new, &e_old_total, &best_total);
    __builtin_amdgcn_wave_barrier();
  }
  gd_slot_store(P, ws, r, acs);
}

/* Wavefront 0: pose buffer pb from movematrix m and the conformer staged
 * at conf (lane-wise atoms; the movematrix and centre by lane 0). */
__device__ __forceinline__ void gd_pose_put(const gd_pack &P, gd_lds_f32 *pb, const float m[6],
                                            const gd_lds_f32 *conf)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const float bx0 = conf[0 * GD_WAVE + lane], by0 = conf[1 * GD_WAVE + lane];
  const float bz0 = conf[2 * GD_WAVE + lane], bx1 = conf[3 * GD_WAVE + lane];
  const float by1 = conf[4 * GD_WAVE + lane], bz1 = conf[5 * GD_WAVE + lane];
  float ax0, ay0, az0, ax1, ay1, az1, cen[3];
  gd_place(P, m, bx0, by0, bz0, bx1, by1, bz1, &ax0, &ay0, &az0, &ax1, &ay1, &az1, cen);
  pb[0 * GD_WAVE + lane] = ax0;
  pb[1 * GD_WAVE + lane] = ay0;
  pb[2 * GD_WAVE + lane] = az0;
  pb[3 * GD_WAVE + lane] = ax1;
  pb[4 * GD_WAVE + lane] = ay1;
  pb[5 * GD_WAVE + lane] = az1;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k)
      pb[GD_PB_M + k] = m[k];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      pb[GD_PB_CEN + k] = cen[k];
  }
}

/* Items of the pose in buffer pb (split launches stage the extras). */
__device__ __forceinline__ void gd_pose_items(const gd_pack &P, const char *smem,
                                              gd_lds_f32 *A, const gd_lds_f32 *pb, int prt,
                                              const gd_split_sizes &Z)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  float xyz[6];
#pragma unroll
  for (int k = 0; k < 6; ++k)
    xyz[k] = pb[k * GD_WAVE + lane];
  gd_energy_items<true>(P, smem, A, prt, Z, xyz[0], xyz[1], xyz[2], xyz[3], xyz[4], xyz[5]);
}

/* n_steps MC steps from step s0 of replica r on the SPLIT wavefronts of the
 * workgroup, `role` = this wavefront's index.  Per step: wavefront 0 names
 * the pose buffer of the step and clears the item counter; barrier; every
 * wavefront takes items -- wavefront 0 only after placing, in the two other
 * buffers, both poses step s + 1 may start from (the move from the current
 * pose if this one is rejected, from the proposed one if it is accepted:
 * the same draws and arithmetic as placing it after the decision); barrier;
 * the sums; barrier; wavefront 0 adds the terms up, decides, and takes the
 * buffer that the decision selects.  Wavefront 0 keeps the state, the
 * decisions and the records.  One code path for every role (the item and
 * sum code is shared, not repeated per role). */
template <int SPLIT>
__device__ __forceinline__ void gd_replica_cycle_split(const gd_pack &P, const char *smem,
                                                       gd_lds_f32 *ws, gd_lds_f32 *A,
                                                       int role, int r, int prt, int s0,
                                                       int n_steps)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  gd_lds_f32 *pose = A + GD_SA_POSE;
  const gd_split_sizes Z = gd_split_sizes_of(P);
  const bool lead = role == 0;

  const gd_state S = P.S;
  float e_old_total = 0.0f, best_total = 0.0f, mb = 0.0f;
  int itmp = 0, acs = 0, cur = 0;
  gd_lds_f32 *conf = A + GD_SA_CONF;
  float d[6], u_acc[2] = {0.0f, 0.0f};
  gd_final_consts F;
  const uint32_t seed = P.seed, cid = P.cid;
  if (lead) {
    gd_slot_load(P, ws, r);
    __builtin_amdgcn_wave_barrier();
    e_old_total = S.e[r * 10 + GD_E_TOTAL];
    best_total = S.e_best[r * 10 + GD_E_TOTAL];
    itmp = S.idx_tmp[r];
    mb = P.minus_beta[itmp];
    F = gd_final_consts_of(P);
    {
      float bx0, by0, bz0, bx1, by1, bz1;
      gd_load_conformer(P, gd_uload(P.rep_lig + r), &bx0, &by0, &bz0, &bx1, &by1, &bz1);
      conf[0 * GD_WAVE + lane] = bx0;
      conf[1 * GD_WAVE + lane] = by0;
      conf[2 * GD_WAVE + lane] = bz0;
      conf[3 * GD_WAVE + lane] = bx1;
      conf[4 * GD_WAVE + lane] = by1;
      conf[5 * GD_WAVE + lane] = bz1;
    }
    /* the pose of the first step, in buffer 0 */
    float m[6];
    gd_draws(P, r, s0, seed, cid, d, u_acc);
#pragma unroll
    for (int k = 0; k < 6; ++k)
      m[k] = d[k] + ws[GD_WS_MOLD + k];
    gd_pose_put(P, pose, m, conf);
    __builtin_amdgcn_wave_barrier();
  }

  for (int s = s0; s < s0 + n_steps; ++s) {
    if (lead) {
      const gd_lds_f32 *pc = pose + cur * GD_PB_FLOATS;
      if (P.rec != nullptr && s % P.rec_stride == 0)
        gd_record(P, ws, r, itmp, s);
      if (lane == 0) {
        *(gd_lds_i32 *) (A + GD_SA_CTR) = 0;
        *(gd_lds_i32 *) (A + GD_SA_SEL) = cur;
#pragma unroll
        for (int k = 0; k < 6; ++k)
          ws[GD_WS_MNEW + k] = pc[GD_PB_M + k];
      }
    }
    __syncthreads();
    const int sel = __builtin_amdgcn_readfirstlane(*(const gd_lds_i32 *) (A + GD_SA_SEL));
    const gd_lds_f32 *pc = pose + sel * GD_PB_FLOATS;

    /* wavefront 0: both poses of step s + 1, buffer ia if this move is
     * rejected, ib if it is accepted */
    const int ia = sel == 2 ? 0 : sel + 1;
    const int ib = ia == 2 ? 0 : ia + 1;
    float u_next[2] = {0.0f, 0.0f}, edst = 0.0f;
    if (lead) {
      if (s + 1 < s0 + n_steps) {
        gd_draws(P, r, s + 1, seed, cid, d, u_next);
        float ma[6], mbv[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          ma[k] = d[k] + ws[GD_WS_MOLD + k];
          mbv[k] = d[k] + pc[GD_PB_M + k];
        }
        gd_pose_put(P, pose + ia * GD_PB_FLOATS, ma, conf);
        gd_pose_put(P, pose + ib * GD_PB_FLOATS, mbv, conf);
      }
      float cen[3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
        cen[k] = pc[GD_PB_CEN + k];
      edst = gd_edst(P, cen);
    }
    gd_pose_items(P, smem, A, pc, prt, Z);
    __syncthreads();
    gd_energy_sums(A, role, SPLIT, Z);
    __syncthreads();
    if (lead) {
      float e_new[GD_E_N], cm_new;
      gd_energy_combine(P, A, F, edst, e_new, &cm_new);
      const int took = gd_accept(P, ws, r, s, u_acc, mb, e_new, cm_new, &e_old_total,
                                 &best_total);
      acs += took;
      cur = took ? ib : ia;
      u_acc[0] = u_next[0];
      u_acc[1] = u_next[1];
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (lead)
    gd_slot_store(P, ws, r, acs);
}

/* ------------------------------------------------------------------------ */
/* Temperature exchange (exchangereplicas_d.cu:28-78), one wavefront per     */
/* chain.  The pairs of one sweep are disjoint in temperature, so each       */
/* pair's decision reads only the two slots it may swap: lane j decides pair */
/* j, with the draws and arithmetic of the serial walk, and writes the two   */
/* slots' new temperature indices and acceptance counts.  Then the step      */
/* counters advance and the work queues are re-armed: by the one workgroup   */
/* of a launch over few chains, else by the last workgroup to finish.        */
/* ------------------------------------------------------------------------ */
#define GD_MAXTMP_DEV 64
#define GD_XWAVES 4        /* wavefronts per workgroup, many-chain form */
#define GD_X1_WAVES 16     /* wavefronts of the one-workgroup form */
#define GD_X1_MAX_CHAINS 32
/* LDS words one exchanging wavefront uses */
#define GD_X_WORDS (2 * GD_MAXTMP_DEV)

__device__ __forceinline__ void gd_advance_all(const gd_pack *__restrict__ packs, int n_cx,
                                               int *queue, int n_queue)
{
  const int n = n_cx > n_queue ? n_cx : n_queue;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (i < n_queue)
      queue[i] = 0;
    if (i < n_cx) {
      const gd_pack &P = packs[i];
      const int s0 = P.step[0];
      const int k = gd_cycle_steps(P, s0);
      if (k > 0) {
        P.step[0] = s0 + k;
        P.step[1] = P.step[1] + 1;
      }
    }
  }
}

/* Chain g (none if g >= n_chain) by this wavefront, LDS scratch sx of
 * GD_X_WORDS words; every wavefront of the workgroup calls it (it holds two
 * workgroup barriers). */
__device__ __forceinline__ void gd_exchange_wave(const gd_pack *__restrict__ packs,
                                                 const gd_chain *__restrict__ chains,
                                                 int n_chain, int g, int *sx)
{
  int *s_orders = sx;                                 /* temperature -> slot */
  float *s_ener = (float *) (sx + GD_MAXTMP_DEV);     /* slot -> total energy */
  const int lane = threadIdx.x & (GD_WAVE - 1);
  int active = 0, n_tmp = 0, n_lig = 0, base = 0, x = 0, c = 0;
  const gd_pack *P = nullptr;
  if (g < n_chain) {
    const gd_chain ch = chains[g];
    P = &packs[ch.cx];
    n_tmp = P->n_tmp;
    active = P->exchange && n_tmp >= 2 && gd_cycle_steps(*P, P->step[0]) > 0;
    if (active) {
      x = P->step[1];
      c = ch.c;
      n_lig = P->n_lig;
      const int p = c / n_lig;
      base = n_tmp * n_lig * p + (c - p * n_lig);
      if (lane < n_tmp) {
        const int a = base + n_lig * lane;
        s_ener[lane] = P->S.e[a * 10 + GD_E_TOTAL];
        s_orders[P->S.idx_tmp[a]] = lane;
      }
    }
  }
  __syncthreads();
  if (active) {
    /* cycle x is launch x % lpw of its steps_per_dump window; the mode
     * alternates from 0 in every window (gpusingle.cu:14-18) */
    const int mode = (x % P->lpw) & 1;
    const int maxpair = n_tmp / 2 - ((n_tmp % 2 == 0) && mode == 1 ? 1 : 0);
    const int j = lane;
    if (j < maxpair) {
      const gd_u32x4 u4 = gd_philox4x32_10((uint32_t) c, (uint32_t) x, GD_CTR_EXCHANGE,
                                           (uint32_t) (j >> 2), P->seed, P->cid);
      const int q = j & 3;
      const uint32_t uw = q == 0 ? u4.v[0] : q == 1 ? u4.v[1] : q == 2 ? u4.v[2] : u4.v[3];
      const float u = gd_exchange_draw(P->fixed_draws, uw);
      const int i1 = 2 * j + mode;
      const int i2 = i1 + 1;
      const int o1 = s_orders[i1];
      const int o2 = s_orders[i2];
      const float delta = (P->minus_beta[i1] - P->minus_beta[i2]) * (s_ener[o2] - s_ener[o1]);
      if (gd_exp(delta) > u) {
        const gd_state S = P->S;
        const int a1 = base + n_lig * o1;
        const int a2 = base + n_lig * o2;
        S.idx_tmp[a1] = i2;
        S.idx_tmp[a2] = i1;
        S.acs_x[a1] += 1;
        S.acs_x[a2] += 1;
      }
    }
  }
  __syncthreads();
}

/* Exchange of at most GD_X1_MAX_CHAINS chains by one workgroup, its
 * wavefronts taking chains in turn, then the advance. */
__global__ void __launch_bounds__(GD_WAVE * GD_X1_WAVES)
gd_exchange1_kernel(const gd_pack *__restrict__ packs, const gd_chain *__restrict__ chains,
                    int n_chain, int n_cx, int *queue, int n_queue)
{
  __shared__ int s_x[GD_X1_WAVES * GD_X_WORDS];
  const int nw = (int) (blockDim.x / GD_WAVE);
  const int wv = threadIdx.x / GD_WAVE;
  for (int g0 = 0; g0 < n_chain; g0 += nw)
    gd_exchange_wave(packs, chains, n_chain, g0 + wv, s_x + wv * GD_X_WORDS);
  gd_advance_all(packs, n_cx, queue, n_queue);
}

/* Exchange of chains GD_XWAVES per workgroup; the last workgroup to finish
 * advances.  Its count orders only the other workgroups' reads of the step
 * counters (all consumed before they count) before the advance's writes:
 * nothing else written here is read here, so no fence is taken. */
__global__ void __launch_bounds__(GD_WAVE * GD_XWAVES)
gd_exchange_kernel(const gd_pack *__restrict__ packs, const gd_chain *__restrict__ chains,
                   int n_chain, int n_cx, int *queue, int n_queue, int *done)
{
  __shared__ int s_x[GD_XWAVES * GD_X_WORDS];
  __shared__ int s_last;
  const int wv = threadIdx.x / GD_WAVE;
  gd_exchange_wave(packs, chains, n_chain, blockIdx.x * GD_XWAVES + wv, s_x + wv * GD_X_WORDS);
  if (threadIdx.x == 0)
    s_last = atomicAdd(done, 1) == (int) gridDim.x - 1;
  __syncthreads();
  if (s_last) {
    gd_advance_all(packs, n_cx, queue, n_queue);
    if (threadIdx.x == 0)
      atomicExch(done, 0);
  }
}

/* After MC (no temperature chains to exchange): step[0] += steps of this
 * cycle, step[1] += 1; threads 0 .. n_queue-1 re-arm the persistent work
 * queues. */
__global__ void __launch_bounds__(256)
gd_advance_kernel(const gd_pack *__restrict__ packs, int n_cx, int *queue, int n_queue)
{
  if (blockIdx.x == 0)
    gd_advance_all(packs, n_cx, queue, n_queue);
}

/* Issue priority of this wavefront in its SIMD (s_setprio takes an
 * immediate): with launches of several streams sharing the device, the
 * higher one issues first whenever both have an instruction ready.  It
 * changes no result. */
__device__ __forceinline__ void gd_set_prio(int p)
{
  if (p == 1)
    __builtin_amdgcn_s_setprio(1);
  else if (p == 2)
    __builtin_amdgcn_s_setprio(2);
  else if (p >= 3)
    __builtin_amdgcn_s_setprio(3);
}

/* ------------------------------------------------------------------------ */
/* MC kernel                                                                 */
/* mode 0: one cycle (steps from each complex's device counter)              */
/* mode 1: initial energies (44.5 * move_scale displacement if P.away)       */
/* queue != nullptr: persistent -- tasks are taken from *queue until it      */
/* passes n_tasks (every workgroup exits once the list is drained).          */
/* ------------------------------------------------------------------------ */
/* WPG: waves (replicas) per workgroup.  MINW: waves per SIMD the register
 * allocation must allow (1 = the compiler's choice, about 90 VGPRs, 5 waves;
 * 6 = 80 VGPRs; 8 = 64 VGPRs with a few spill slots) -- the launch shape
 * and this are picked by timing. */
template <int WPG, int MINW>
__global__ void __launch_bounds__(GD_WAVE *WPG) __attribute__((amdgpu_waves_per_eu(MINW, 8)))
gd_mc_kernel(const gd_pack *__restrict__ packs, const gd_task *__restrict__ tasks,
             int n_tasks, int *queue, int mode, int wslot_off, int prio)
{
  extern __shared__ __attribute__((aligned(16))) char smem[];
  gd_set_prio(prio);
  __shared__ int s_task;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / GD_WAVE);
  gd_lds_f32 *ws = (gd_lds_f32 *) (smem + wslot_off) + w * GD_WSLOT_FLOATS;
  int staged_cx = -1, staged_prt = -1;
  int t = queue ? -1 : (int) blockIdx.x;

  for (;;) {
    if (queue) {
      if (threadIdx.x == 0)
        s_task = atomicAdd(queue, 1);
      __syncthreads();
      /* the task index is workgroup-uniform: keep it (and every pack field
       * read through it) in scalar registers */
      t = __builtin_amdgcn_readfirstlane(s_task);
      __syncthreads();
    } else if (t != (int) blockIdx.x) {
      break;  /* grid form: one task per workgroup */
    }
    if (t >= n_tasks)
      break;
    const gd_task task = tasks[t];
    const gd_pack &P = packs[task.cx];
    int s0 = 0, n_steps = 0;
    if (mode == 0) {
      s0 = P.step[0];
      n_steps = gd_cycle_steps(P, s0);
    }
    if (mode != 0 || n_steps > 0) {
      if (task.cx != staged_cx || task.prt != staged_prt) {
        __syncthreads();
        gd_stage(P, task.prt, smem);
        __syncthreads();
        staged_cx = task.cx;
        staged_prt = task.prt;
      }
      if (w < task.r_count) {
        const int r = task.r_begin + w;
        if (mode == 0)
          gd_replica_cycle(P, smem, ws, r, task.prt, s0, n_steps);
        else
          gd_replica_init(P, smem, r);
      }
    }
    if (!queue)
      t = -2;
  }
}

/* MC cycle with SPLIT wavefronts per replica and one replica per workgroup
 * (tasks of one replica), grid form or persistent as gd_mc_kernel: for
 * launches with fewer replicas than the device has SIMDs, where one
 * wavefront per replica leaves most of the device idle and every step
 * waits on its own latency. */
template <int SPLIT>
__global__ void __launch_bounds__(GD_WAVE *SPLIT)
gd_mc_split_kernel(const gd_pack *__restrict__ packs, const gd_task *__restrict__ tasks,
                   int n_tasks, int *queue, int wslot_off, int prio)
{
  extern __shared__ __attribute__((aligned(16))) char smem[];
  gd_set_prio(prio);
  __shared__ int s_task;
  const int role = __builtin_amdgcn_readfirstlane(threadIdx.x / GD_WAVE);
  gd_lds_f32 *ws = (gd_lds_f32 *) (smem + wslot_off);
  gd_lds_f32 *A = ws + GD_WSLOT_FLOATS;
  int staged_cx = -1, staged_prt = -1;
  int t = queue ? -1 : (int) blockIdx.x;

  for (;;) {
    if (queue) {
      if (threadIdx.x == 0)
        s_task = atomicAdd(queue, 1);
      __syncthreads();
      t = __builtin_amdgcn_readfirstlane(s_task);
      __syncthreads();
    } else if (t != (int) blockIdx.x) {
      break;
    }
    if (t >= n_tasks)
      break;
    const gd_task task = tasks[t];
    const gd_pack &P = packs[task.cx];
    const int s0 = P.step[0];
    const int n_steps = gd_cycle_steps(P, s0);
    if (n_steps > 0) {
      if (task.cx != staged_cx || task.prt != staged_prt) {
        __syncthreads();
        gd_stage(P, task.prt, smem);
        gd_stage_split_extra(P, task.prt, smem);
        __syncthreads();
        staged_cx = task.cx;
        staged_prt = task.prt;
      }
      gd_replica_cycle_split<SPLIT>(P, smem, ws, A, role, task.r_begin, task.prt, s0,
                                    n_steps);
    }
    if (!queue)
      t = -2;
  }
}


/* ------------------------------------------------------------------------ */
/* Energies of given poses (FP32).                                           */
/* ------------------------------------------------------------------------ */
__global__ void __launch_bounds__(GD_WAVE * 8)
gd_score_kernel(const gd_pack *__restrict__ packs, const gd_task *__restrict__ tasks,
                const int *__restrict__ pose_ids, const int *__restrict__ lig_conf,
                const float *__restrict__ move, float *__restrict__ energy)
{
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const gd_task task = tasks[blockIdx.x];
  const gd_pack &P = packs[task.cx];
  gd_stage(P, task.prt, smem);
  __syncthreads();
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / GD_WAVE);
  if (w >= task.r_count)
    return;
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const int i = pose_ids[task.r_begin + w];
  float bx0, by0, bz0, bx1, by1, bz1;
  gd_load_conformer(P, lig_conf[i], &bx0, &by0, &bz0, &bx1, &by1, &bz1);
  float m[6];
#pragma unroll
  for (int k = 0; k < 6; ++k)
    m[k] = move[i * 6 + k];
  float ax0, ay0, az0, ax1, ay1, az1, cen[3], e[GD_E_N];
  gd_place(P, m, bx0, by0, bz0, bx1, by1, bz1, &ax0, &ay0, &az0, &ax1, &ay1, &az1, cen);
  float cm;
  gd_energy_wave(P, smem, task.prt, ax0, ay0, az0, ax1, ay1, az1, cen, e, &cm);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < GD_E_N; ++k)
      energy[i * GD_E_N + k] = e[k];
  }
}

/* ------------------------------------------------------------------------ */
/* FP64 re-score: the reference formulas (calcenergy_d.cu:69-309) in double, */
/* one wavefront per pose, reading the pack straight from HBM.              */
/* ------------------------------------------------------------------------ */
/* The Bayesian total in FP64: the same fitted pdfs and the same floor (1e-8
 * for a zero or non-finite pdf or an exponential factor below e^-86). */
__device__ double gd_pdf64(int family, double x, double loc, double scale)
{
  const double pi = 3.14159265358979323846;
  double v, arg = 0.0;
  if (family == 0) {
    const double d = x - loc;
    arg = -d * d / (2.0 * scale * scale);
    v = exp(arg) / (scale * sqrt(2.0 * pi));
  } else if (family == 1) {
    const double t = (x - loc) / scale;
    v = 1.0 / (pi * scale * (1.0 + t * t));
  } else if (family == 2) {
    arg = -fabs((x - loc) / scale);
    const double e = exp(arg);
    v = e / ((1.0 + e) * (1.0 + e)) / scale;
  } else {
    const double nx = (x - loc) / scale;
    if (!(nx > 0.0))
      return 1e-8;
    arg = -(nx - 1.0) * (nx - 1.0) / (2.0 * nx);
    v = exp(arg) / (sqrt(2.0 * pi * nx * nx * nx) * scale);
  }
  return (v > 0.0 && v < HUGE_VAL && arg >= -86.0) ? v : 1e-8;
}

__device__ double gd_bayes_total64(const double nf[9])
{
  const float hi[9][2] = GD_BAYES_HIGH_INIT;
  const float lo[9][2] = GD_BAYES_LOW_INIT;
  const int fam[9] = GD_BAYES_FAMILY_INIT;
  double ph = 0.0, pl = 0.0;
  for (int k = 0; k < 9; ++k) {
    ph += log10(gd_pdf64(fam[k], nf[k], hi[k][0], hi[k][1]));
    pl += log10(gd_pdf64(fam[k], nf[k], lo[k][0], lo[k][1]));
  }
  return pl - ph;
}

__device__ void gd_energy64_wave(const gd_pack &P, int lconf, int pconf, const float *move,
                                 double *__restrict__ out)
{
  const int lane = threadIdx.x & (GD_WAVE - 1);
  const int lna = P.lna, pnp = P.pnp;
  double m[6];
  for (int k = 0; k < 6; ++k)
    m[k] = (double) move[k];
  const double s1 = sin(m[3]), c1 = cos(m[3]);
  const double s2 = sin(m[4]), c2 = cos(m[4]);
  const double s3 = sin(m[5]), c3 = cos(m[5]);
  const double R[9] = {c1 * c2, c1 * s2 * s3 - c3 * s1, s1 * s3 + c1 * c3 * s2,
                       c2 * s1, c1 * c3 + s1 * s2 * s3, c3 * s1 * s2 - c1 * s3,
                       -s2,     c2 * s3,                c2 * c3};
  const double cen[3] = {(double) P.center[0] + m[0], (double) P.center[1] + m[1],
                         (double) P.center[2] + m[2]};
  const float4 *cx = P.lig_xyz + (size_t) lconf * lna;
  const float4 *pp = P.pt_pos + (size_t) pconf * P.pnp_pad;
  const float2 *pe = P.pt_eh + (size_t) pconf * P.pnp_pad;
  const double lj0 = P.sc.lj0, lj1 = P.sc.lj1, el0 = P.sc.el0, el1 = P.sc.el1;
  const double a1 = 4.0 - 3.0 * el0, b1 = 2.0 * el0 - 3.0;

  double evdw = 0, eele = 0, epmf = 0, epsp = 0, ehdb = 0, ehpc = 0, ekde = 0, elhm = 0;
  for (int l = 0; l < lna; ++l) {
    const float4 o = cx[l];
    const double lx = R[0] * o.x + R[1] * o.y + R[2] * o.z + cen[0];
    const double ly = R[3] * o.x + R[4] * o.y + R[5] * o.z + cen[1];
    const double lz = R[6] * o.x + R[7] * o.y + R[8] * o.z + cen[2];
    const int row = P.lig_row[l];
    const int lt = P.lig_type[l];
    const double lq = P.lig_q[l];
    const gd_pair_param *trow = P.tab + row * GD_ROW_STRIDE;
    double hpc1 = 0.0;
    for (int p = lane; p < pnp; p += GD_WAVE) {
      const float4 q = pp[p];
      const float2 eh = pe[p];
      const uint32_t bits = __float_as_uint(q.w);
      const gd_pair_param tp = trow[bits >> GD_PT_TYPE_SHIFT];
      const double dx = lx - q.x, dy = ly - q.y, dz = lz - q.z;
      const double d2 = dx * dx + dy * dy + dz * dz;
      const double d = sqrt(d2);
      if (eh.y != 0.0f && d2 <= 81.0) {
        const double u = d2 / 81.0;
        hpc1 += (double) eh.y * (1.0 - (3.5 * u - 4.5 * u * u + 2.5 * u * u * u - 0.5 * u * u * u * u));
      }
      const double p1 = (double) tp.p1a / (d2 * d2 * d2 * d2 * d);
      const double p2 = (double) tp.p2a / (d2 * d2 * d2);
      const double p4 = p1 * lj0 * (1.0 + lj1 * d2) + 1.0;
      evdw += (p1 - p2) / p4;
      const double s = el1 * d;
      const double g = s < 1.0 ? el0 + a1 * s * s + b1 * s * s * s : 1.0 / s;
      eele += lq * (double) eh.x * g;
      const double dm = d - (double) tp.pmf0;
      epmf += (double) tp.pmf1 / (1.0 + exp((-0.5 * d + 6.0) * dm));
      if (dm <= 0.0)
        epsp += (double) P.psp_pt[((size_t) pconf * P.n_rows + row) * P.pnp_pad + p];
      if (tp.hdb0 > 0.1f) {
        const double h = (d - (double) tp.hdb0) * (double) tp.hdb1;
        ehdb += (double) tp.hdb1 * exp(-0.5 * h * h);
      }
    }
    hpc1 = gd_wave_sum64(hpc1);
    const double z = (hpc1 - (double) P.hpl0[lt]) / (double) P.hpl1[lt];
    ehpc += 0.5 * z * z - (double) P.hpl2[lt];

    const int k0 = P.kde_seg[2 * row], k1 = P.kde_seg[2 * row + 1];
    double v = 0.0;
    for (int k = k0 + lane; k < k1; k += GD_WAVE) {
      const float4 q = P.kde[k];
      const double ex = lx - q.x, ey = ly - q.y, ez = lz - q.z;
      v += exp((double) P.kde2 * (ex * ex + ey * ey + ez * ez));
    }
    v = gd_wave_sum64(v);
    if (k1 > k0)
      ekde += v / (double) (k1 - k0);
  }
  evdw = gd_wave_sum64(evdw);
  eele = gd_wave_sum64(eele);
  epmf = gd_wave_sum64(epmf);
  epsp = gd_wave_sum64(epsp);
  ehdb = gd_wave_sum64(ehdb);
  ekde = ekde / (double) P.kde3;

  for (int mm = 0; mm < P.pos; ++mm) {
    double s = 0.0, c = 0.0;
    for (int l = lane; l < lna; l += GD_WAVE) {
      const float4 t = P.mcs[(size_t) mm * lna + l];
      if (t.w != 0.0f) {
        const float4 o = cx[l];
        const double lx = R[0] * o.x + R[1] * o.y + R[2] * o.z + cen[0];
        const double ly = R[3] * o.x + R[4] * o.y + R[5] * o.z + cen[1];
        const double lz = R[6] * o.x + R[7] * o.y + R[8] * o.z + cen[2];
        const double dx = lx - t.x, dy = ly - t.y, dz = lz - t.z;
        s += dx * dx + dy * dy + dz * dz;
        c += 1.0;
      }
    }
    s = gd_wave_sum64(s);
    c = gd_wave_sum64(c);
    if (c > 0.0)
      elhm += (double) P.mcs_tcc[mm] * sqrt(s / c);
  }
  elhm = P.pos > 0 ? log(elhm / (double) P.pos) : 0.0;
  const double ddx = cen[0] - P.pocket_center[0], ddy = cen[1] - P.pocket_center[1],
               ddz = cen[2] - P.pocket_center[2];
  const double edst = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);

  if (lane == 0) {
    const double fl = (double) lna;
    double raw[9];
    raw[GD_F_VDW] = evdw / fl;
    raw[GD_F_ELE] = eele / fl;
    raw[GD_F_PMF] = epmf / fl;
    raw[GD_F_HPC] = ehpc / fl;
    raw[GD_F_HDB] = -ehdb / fl / sqrt(2.0 * 3.14159265358979323846);
    raw[GD_F_DST] = edst;
    raw[GD_F_PSP] = epsp / fl;
    raw[GD_F_KDE] = ekde / fl;
    raw[GD_F_LHM] = elhm;
    const int f2s[9] = GD_FILE_TO_SLOT_INIT;
    double slot[10];
    for (int k = 0; k < 9; ++k)
      slot[f2s[k]] = raw[k];
    double total = 0.0;
    if (P.term_order == 0) {
      for (int k = 0; k < 9; ++k) {
        const double v = (double) P.a_para[k] * raw[k] + (double) P.b_para[k];
        slot[f2s[k]] = v;
        total += (double) P.w[k] * v;
      }
    } else {
      for (int k = 0; k < 9; ++k) {
        const double v = (double) P.a_para[k] * slot[k] + (double) P.b_para[k];
        slot[k] = v;
        total += (double) P.w[k] * v;
      }
    }
    if (P.combine == 1) {
      double nf[9];
      for (int k = 0; k < 9; ++k)
        nf[k] = slot[f2s[k]];
      total = gd_bayes_total64(nf);
    }
    slot[9] = total;
    for (int k = 0; k < 10; ++k)
      out[k] = slot[k];
  }
}


/* FP64 energies of caller-given poses of one complex. */
__global__ void __launch_bounds__(GD_WAVE)
gd_rescore64_kernel(const gd_pack *__restrict__ packs, int cx_index,
                    const int *__restrict__ lig_conf, const int *__restrict__ prt_conf,
                    const float *__restrict__ move, int move_stride, int n_pose,
                    double *__restrict__ energy)
{
  const int i = blockIdx.x;
  if (i >= n_pose)
    return;
  gd_energy64_wave(packs[cx_index], lig_conf[i], prt_conf[i], move + (size_t) i * move_stride,
                   energy + (size_t) i * 10);
}

/* FP64 energies of the best poses of listed replicas (complex, replica) of
 * many complexes, one launch: the read-back of finished jobs. */
__global__ void __launch_bounds__(GD_WAVE)
gd_rescore64_best_kernel(const gd_pack *__restrict__ packs, const int2 *__restrict__ items,
                         int n, double *__restrict__ energy)
{
  const int i = blockIdx.x;
  if (i >= n)
    return;
  const int2 it = items[i];
  const gd_pack &P = packs[it.x];
  gd_energy64_wave(P, P.rep_lig[it.y], P.rep_prt[it.y], P.S.mv_best + (size_t) it.y * 8,
                   energy + (size_t) i * 10);
}

/* ------------------------------------------------------------------------ */
/* host-callable launch wrappers                                             */
/* ------------------------------------------------------------------------ */
template <int WPG, int MINW>
static hipError_t launch_mc_var(const gd_pack *packs, const gd_task *tasks, int n_tasks,
                                int *queue, int grid, int mode, int stage_bytes, int prio,
                                hipStream_t st)
{
  const int lds = stage_bytes + WPG * GD_WSLOT_FLOATS * 4;
  hipLaunchKernelGGL((gd_mc_kernel<WPG, MINW>), dim3(grid), dim3(GD_WAVE * WPG), lds, st,
                     packs, tasks, n_tasks, queue, mode, stage_bytes, prio);
  return hipGetLastError();
}

template <int WPG>
static hipError_t launch_mc_wpg(const gd_launch_cfg &C, const gd_pack *packs,
                                const gd_task *tasks, int n_tasks, int *queue, int grid,
                                int mode, hipStream_t st)
{
  switch (C.occ) {
  case 6:
    return launch_mc_var<WPG, 6>(packs, tasks, n_tasks, queue, grid, mode, C.stage_bytes,
                                 C.prio, st);
  case 8:
    return launch_mc_var<WPG, 8>(packs, tasks, n_tasks, queue, grid, mode, C.stage_bytes,
                                 C.prio, st);
  default:
    return launch_mc_var<WPG, 1>(packs, tasks, n_tasks, queue, grid, mode, C.stage_bytes,
                                 C.prio, st);
  }
}

template <int SPLIT>
static hipError_t launch_split(const gd_pack *packs, const gd_task *tasks, int n_tasks,
                               int *queue, int grid, int stage_bytes, int prio, hipStream_t st)
{
  const int lds = stage_bytes + GD_WSLOT_FLOATS * 4 + (int) GD_SPLIT_AREA_BYTES;
  hipLaunchKernelGGL(gd_mc_split_kernel<SPLIT>, dim3(grid), dim3(GD_WAVE * SPLIT), lds, st,
                     packs, tasks, n_tasks, queue, stage_bytes, prio);
  return hipGetLastError();
}

hipError_t gd_launch_mc(const gd_launch_cfg &C, const gd_pack *packs, const gd_task *tasks,
                        int n_tasks, int *queue, int grid, int mode, hipStream_t st)
{
  if (n_tasks <= 0 || grid <= 0)
    return hipSuccess;
  if (mode == 0 && C.split > 1) {
    switch (C.split) {
    case 4: return launch_split<4>(packs, tasks, n_tasks, queue, grid, C.stage_bytes, C.prio, st);
    case 8: return launch_split<8>(packs, tasks, n_tasks, queue, grid, C.stage_bytes, C.prio, st);
    case 16:
      return launch_split<16>(packs, tasks, n_tasks, queue, grid, C.stage_bytes, C.prio, st);
    default: return hipErrorInvalidValue;
    }
  }
  switch (C.wpg) {
  case 4: return launch_mc_wpg<4>(C, packs, tasks, n_tasks, queue, grid, mode, st);
  case 16: return launch_mc_wpg<16>(C, packs, tasks, n_tasks, queue, grid, mode, st);
  default: return launch_mc_wpg<8>(C, packs, tasks, n_tasks, queue, grid, mode, st);
  }
}

hipError_t gd_launch_exchange(const gd_pack *packs, const gd_chain *chains, int n_chain,
                              int n_cx, int *queue, int n_queue, int *done, hipStream_t st)
{
  if (n_chain <= 0)
    return gd_launch_advance(packs, n_cx, queue, n_queue, st);
  if (n_chain <= GD_X1_MAX_CHAINS) {
    const int nw = n_chain < GD_X1_WAVES ? n_chain : GD_X1_WAVES;
    hipLaunchKernelGGL(gd_exchange1_kernel, dim3(1), dim3(GD_WAVE * nw), 0, st, packs, chains,
                       n_chain, n_cx, queue, n_queue);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(gd_exchange_kernel, dim3((n_chain + GD_XWAVES - 1) / GD_XWAVES),
                     dim3(GD_WAVE * GD_XWAVES), 0, st, packs, chains, n_chain, n_cx, queue,
                     n_queue, done);
  return hipGetLastError();
}

hipError_t gd_launch_advance(const gd_pack *packs, int n_cx, int *queue, int n_queue,
                             hipStream_t st)
{
  hipLaunchKernelGGL(gd_advance_kernel, dim3(1), dim3(256), 0, st, packs, n_cx, queue,
                     n_queue);
  return hipGetLastError();
}

hipError_t gd_launch_score(const gd_pack *packs, const gd_task *tasks, int n_tasks,
                           const int *pose_ids, const int *lig_conf, const float *move,
                           int stage_bytes, float *energy, hipStream_t st)
{
  if (n_tasks <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(gd_score_kernel, dim3(n_tasks), dim3(GD_WAVE * 8), stage_bytes, st,
                     packs, tasks, pose_ids, lig_conf, move, energy);
  return hipGetLastError();
}

hipError_t gd_launch_rescore64_best(const gd_pack *packs, const int2 *items, int n,
                                    double *energy, hipStream_t st)
{
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(gd_rescore64_best_kernel, dim3(n), dim3(GD_WAVE), 0, st, packs, items, n,
                     energy);
  return hipGetLastError();
}

hipError_t gd_launch_rescore64(const gd_pack *packs, int cx_index, const int *lig_conf,
                               const int *prt_conf, const float *move, int move_stride,
                               int n_pose, double *energy, hipStream_t st)
{
  if (n_pose <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(gd_rescore64_kernel, dim3(n_pose), dim3(GD_WAVE), 0, st, packs,
                     cx_index, lig_conf, prt_conf, move, move_stride, n_pose, energy);
  return hipGetLastError();
}

template <int WPG, int MINW>
static hipError_t set_attr_var(int bytes)
{
  return hipFuncSetAttribute((const void *) gd_mc_kernel<WPG, MINW>,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             bytes + WPG * GD_WSLOT_FLOATS * 4);
}

template <int WPG>
static hipError_t set_attr_wpg(int bytes)
{
  hipError_t e = set_attr_var<WPG, 1>(bytes);
  if (e == hipSuccess) e = set_attr_var<WPG, 6>(bytes);
  if (e == hipSuccess) e = set_attr_var<WPG, 8>(bytes);
  return e;
}

template <int SPLIT>
static hipError_t set_attr_split(int bytes)
{
  /* a complex whose staging and split area exceed the device's LDS never
   * runs split (the engine checks the same sum), so it needs no attribute */
  int dev = 0, max_lds = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e != hipSuccess)
    return e;
  const int want = bytes + GD_WSLOT_FLOATS * 4 + (int) GD_SPLIT_AREA_BYTES;
  if (want > max_lds)
    return hipSuccess;
  return hipFuncSetAttribute((const void *) gd_mc_split_kernel<SPLIT>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, want);
}

hipError_t gd_set_kernel_attrs(int stage_bytes)
{
  hipError_t e = set_attr_split<4>(stage_bytes);
  if (e == hipSuccess) e = set_attr_split<8>(stage_bytes);
  if (e == hipSuccess) e = set_attr_split<16>(stage_bytes);
  if (e == hipSuccess) e = set_attr_wpg<4>(stage_bytes);
  if (e == hipSuccess) e = set_attr_wpg<8>(stage_bytes);
  if (e == hipSuccess) e = set_attr_wpg<16>(stage_bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *) gd_score_kernel,
                            hipFuncAttributeMaxDynamicSharedMemorySize, stage_bytes);
  return e;
}

template <int WPG, int MINW>
static hipError_t occ_var(int stage_bytes, int *blocks)
{
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocks, gd_mc_kernel<WPG, MINW>, GD_WAVE * WPG, stage_bytes + WPG * GD_WSLOT_FLOATS * 4);
}

template <int WPG>
static hipError_t occ_wpg(int occ, int stage_bytes, int *blocks)
{
  switch (occ) {
  case 6: return occ_var<WPG, 6>(stage_bytes, blocks);
  case 8: return occ_var<WPG, 8>(stage_bytes, blocks);
  default: return occ_var<WPG, 1>(stage_bytes, blocks);
  }
}

template <int SPLIT>
static hipError_t occ_split(int stage_bytes, int *blocks)
{
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocks, gd_mc_split_kernel<SPLIT>, GD_WAVE * SPLIT,
      stage_bytes + GD_WSLOT_FLOATS * 4 + (int) GD_SPLIT_AREA_BYTES);
}

hipError_t gd_mc_occupancy(const gd_launch_cfg &C, int stage_bytes, int *blocks_per_cu)
{
  switch (C.split) {
  case 4: return occ_split<4>(stage_bytes, blocks_per_cu);
  case 8: return occ_split<8>(stage_bytes, blocks_per_cu);
  case 16: return occ_split<16>(stage_bytes, blocks_per_cu);
  default: break;
  }
  switch (C.wpg) {
  case 4: return occ_wpg<4>(C.occ, stage_bytes, blocks_per_cu);
  case 16: return occ_wpg<16>(C.occ, stage_bytes, blocks_per_cu);
  default: return occ_wpg<8>(C.occ, stage_bytes, blocks_per_cu);
  }
}

/* ------------------------------------------------------------------------ */
/* Static device workspace of the allocation-free pose scoring entry point:  */
/* one array per GPU, for the whole process.  Host code never uses the      */
/* array's name as an address: gd_scratch_region resolves each device's     */
/* own copy once, with that device current, and hands out that address.     */
/* ------------------------------------------------------------------------ */
static __device__ __align__(256) char scratch[GD_SCRATCH_BYTES];

#define GD_SCRATCH_MAX_DEVICES 64

hipError_t gd_scratch_region(int device, char **ptr)
{
  static std::mutex mu;
  static char *dev_ptr[GD_SCRATCH_MAX_DEVICES] = {};
  if (device < 0 || device >= GD_SCRATCH_MAX_DEVICES)
    return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!dev_ptr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess)
      return e;
    if (cur != device)
      return hipErrorInvalidDevice;   /* the caller makes `device` current */
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return e;
    if (!p)
      return hipErrorInvalidSymbol;
    dev_ptr[device] = (char *) p;
  }
  *ptr = dev_ptr[device];
  return hipSuccess;
}

/* ------------------------------------------------------------------------ */
/* VALU issue-rate probe: eight independent FMA chains per lane, every       */
/* operand in a vector register (v_fmac_f32_e32), 16 x iters rounds; one     */
/* store per lane keeps the chains live.  Calibrates the FP32 VALU ceiling   */
/* that the MC kernel's instruction rate is read against.                    */
/* ------------------------------------------------------------------------ */
__global__ void __launch_bounds__(256) gd_valu_probe_kernel(float *out, int iters)
{
  const float a = 0.999f + (float) threadIdx.x * 1e-7f;
  const float b = 1e-3f + (float) threadIdx.x * 1e-7f;
  float x[8];
#pragma unroll
  for (int c = 0; c < 8; ++c)
    x[c] = (float) threadIdx.x * 1e-3f + (float) c;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
#pragma unroll
      for (int c = 0; c < 8; ++c)
        x[c] = gd_fma(a, b, x[c]);
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    s += x[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

hipError_t gd_launch_valu_probe(float *out, int blocks, int iters, hipStream_t st)
{
  hipLaunchKernelGGL(gd_valu_probe_kernel, dim3(blocks), dim3(256), 0, st, out, iters);
  return hipGetLastError();
}

/* Issue rate of one VALU instruction kind: eight independent register
 * chains per lane, 16 rounds of them per loop iteration, nothing but
 * registers inside the loop and one store per lane at the end.  The kinds
 * are the ones the energy loops are made of (gd_valu_kind_names). */
#define GD_K8(...)                                                              \
  do {                                                                          \
    _Pragma("unroll") for (int k_ = 0; k_ < 16; ++k_) {                        \
      GD_K1(x0, __VA_ARGS__); GD_K1(x1, __VA_ARGS__); GD_K1(x2, __VA_ARGS__);  \
      GD_K1(x3, __VA_ARGS__); GD_K1(x4, __VA_ARGS__); GD_K1(x5, __VA_ARGS__);  \
      GD_K1(x6, __VA_ARGS__); GD_K1(x7, __VA_ARGS__);                          \
    }                                                                           \
  } while (0)
template <int KIND>
__global__ void __launch_bounds__(256)
    gd_valu_kind_kernel(float *out, int iters, unsigned long long mask, float sv)
{
  float x0 = (float) threadIdx.x * 1e-3f, x1 = x0 + 1.0f, x2 = x0 + 2.0f, x3 = x0 + 3.0f;
  float x4 = x0 + 4.0f, x5 = x0 + 5.0f, x6 = x0 + 6.0f, x7 = x0 + 7.0f;
  const float a = 0.999f, b = 1e-3f;
  int sink = 0;
  for (int i = 0; i < iters; ++i) {
    if (KIND == 0) {
#define GD_K1(x, ...) asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(x) : "v"(a), "v"(b))
      GD_K8(0);
#undef GD_K1
    } else if (KIND == 1) {
#define GD_K1(x, ...) asm volatile("v_fma_f32 %0, %1, %2, %0" : "+v"(x) : "s"(sv), "v"(b))
      GD_K8(0);
#undef GD_K1
    } else if (KIND == 2) {
#define GD_K1(x, ...) asm volatile("v_mul_f32 %0, %0, %1" : "+v"(x) : "v"(a))
      GD_K8(0);
#undef GD_K1
    } else if (KIND == 3) {
#define GD_K1(x, ...) asm volatile("v_max_f32 %0, %0, %1" : "+v"(x) : "v"(b))
      GD_K8(0);
#undef GD_K1
    } else if (KIND == 4) {
#define GD_K1(x, ...) asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(x) : "v"(a), "s"(mask))
      GD_K8(0);
#undef GD_K1
    } else if (KIND == 5) {
#define GD_K1(x, ...) asm volatile("v_lshl_add_u32 %0, %0, 1, %1" : "+v"(x) : "v"(b))
      GD_K8(0);
#undef GD_K1
    } else if (KIND == 6) {
#define GD_K1(x, ...) asm volatile("v_cmp_gt_f32 vcc, %0, %1" : : "v"(x), "v"(a) : "vcc")
      GD_K8(0);
#undef GD_K1
    } else if (KIND == 7) {
#define GD_K1(x, ...) \
  asm volatile("v_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(x))
      GD_K8(0);
#undef GD_K1
    } else if (KIND == 8) {
#define GD_K1(x, ...)                                                    \
  do {                                                                   \
    int s_;                                                              \
    asm volatile("v_readlane_b32 %0, %1, 5" : "=s"(s_) : "v"(x));        \
    sink ^= s_;                                                          \
  } while (0)
      GD_K8(0);
#undef GD_K1
    } else {
      /* FMA and max alternating: how much of a slower kind hides behind
       * the full-rate one */
#define GD_K1(x, ...) asm volatile("v_fmac_f32 %0, %1, %2\n\tv_max_f32 %0, %0, %1" : "+v"(x) : "v"(a), "v"(b))
      GD_K8(0);
#undef GD_K1
    }
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] =
      ((x0 + x1) + (x2 + x3)) + ((x4 + x5) + (x6 + x7)) + (float) sink;
}
#undef GD_K8

hipError_t gd_launch_valu_kind(int kind, float *out, int blocks, int iters, hipStream_t st)
{
  const unsigned long long m = 0x5555555555555555ull;
  switch (kind) {
  case 0: hipLaunchKernelGGL(gd_valu_kind_kernel<0>, dim3(blocks), dim3(256), 0, st, out, iters, m, 1.0f); break;
  case 1: hipLaunchKernelGGL(gd_valu_kind_kernel<1>, dim3(blocks), dim3(256), 0, st, out, iters, m, 1.0f); break;
  case 2: hipLaunchKernelGGL(gd_valu_kind_kernel<2>, dim3(blocks), dim3(256), 0, st, out, iters, m, 1.0f); break;
  case 3: hipLaunchKernelGGL(gd_valu_kind_kernel<3>, dim3(blocks), dim3(256), 0, st, out, iters, m, 1.0f); break;
  case 4: hipLaunchKernelGGL(gd_valu_kind_kernel<4>, dim3(blocks), dim3(256), 0, st, out, iters, m, 1.0f); break;
  case 5: hipLaunchKernelGGL(gd_valu_kind_kernel<5>, dim3(blocks), dim3(256), 0, st, out, iters, m, 1.0f); break;
  case 6: hipLaunchKernelGGL(gd_valu_kind_kernel<6>, dim3(blocks), dim3(256), 0, st, out, iters, m, 1.0f); break;
  case 7: hipLaunchKernelGGL(gd_valu_kind_kernel<7>, dim3(blocks), dim3(256), 0, st, out, iters, m, 1.0f); break;
  case 8: hipLaunchKernelGGL(gd_valu_kind_kernel<8>, dim3(blocks), dim3(256), 0, st, out, iters, m, 1.0f); break;
  default: hipLaunchKernelGGL(gd_valu_kind_kernel<9>, dim3(blocks), dim3(256), 0, st, out, iters, m, 1.0f); break;
  }
  return hipGetLastError();
}
