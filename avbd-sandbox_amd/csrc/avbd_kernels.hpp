// avbd_kernels.hpp — the kernels of one AVBD step on gfx950.
//
// Step pipeline (one HIP stream):
//   broad phase   bp_cells -> bp_scatter -> bp_small/bp_large (count, write)
//   narrow phase  narrow (one thread per old manifold or new pair) + joint_init
//   topology      live compaction, per-body force CSR (adj_*), colouring
//                 rounds to the greedy fixed point, colour buckets, per-body
//                 contact work lists (item_*)
//   solve         body_init, then per iteration: primal per colour, dual,
//                 and velocity after the last main iteration
// References are file:line in the reference repository.
#pragma once
#include <type_traits>

#include "avbd_device.hpp"

namespace avbd {

struct StepParams {
  double dt, gravity, betaLin, betaAng, alpha, gamma, pmin, pmax, margin, stick;
  int iterations, postStab, rotated;
  // the GPU contact model (shaders.js:1105-1367): on or off, the impact speed
  // below which no restitution target is set, the warm-start persistence
  // distance (solver.js:155-162)
  int contactModel;
  double restThreshold, persistence;
};

// ------------------------------------------------------------------ utils
__global__ void k_or_flag(int *p, int bits) { atomicOr(p, bits); }

__global__ void k_fill(int *p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Block-level exclusive scan of 1024 ints (256 threads x 4). With nDev, the
// count is min(n, *nDev) (a count produced on the device; the grid covers n).
__global__ void k_scan_block(const int *in, int *out, int n, int *blockSums, const int *nDev) {
  __shared__ int sh[256];
  if (nDev) n = min(n, *nDev);
  const int base = blockIdx.x * 1024 + threadIdx.x * 4;
  int v[4], s = 0;
  for (int k = 0; k < 4; k++) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - s;
  for (int k = 0; k < 4; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) blockSums[blockIdx.x] = sh[255];
}

// Second and last pass: each 1024-element block adds the sum of the blocks
// before it, which its 256 threads add up themselves from the block sums (a
// few hundred at most for the sizes here, so no third pass is needed); the
// last block also writes the total.
__global__ void k_scan_finish(int *out, int n, const int *blockSums, int nb, int *total, const int *nDev) {
  __shared__ int red[256];
  if (nDev) n = min(n, *nDev);
  const int b = blockIdx.x;
  int s = 0;
  for (int j = threadIdx.x; j < b; j += 256) s += blockSums[j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  const int add = red[0];
  if (b == nb - 1 && threadIdx.x == 0) *total = add + blockSums[b];
  if (add == 0) return;
  const int i = b * 1024 + threadIdx.x * 4;
  for (int k = 0; k < 4; k++)
    if (i + k < n) out[i + k] += add;
}

// ------------------------------------------------------------- broad phase
// Uniform grid over small bodies (radius <= 8x the scene's mean), cell edge h
// >= 2 x the largest small radius, hashed by (scene, cell). Every pair whose
// bounding spheres overlap has centres in neighbouring cells, so the 27-cell
// query returns a superset of the brute-force pairs (broadphase.js:236-243)
// and the exact predicate of solver.js:316-329 selects them.
__device__ __forceinline__ uint32_t cellHash(int s, int x, int y, int z) {
  uint32_t h = uint32_t(x) * 73856093u ^ uint32_t(y) * 19349663u ^ uint32_t(z) * 83492791u ^ uint32_t(s) * 2654435761u;
  h ^= h >> 15;
  h *= 0x2c1b3c6du;
  h ^= h >> 12;
  return h;
}

__global__ void k_bp_cells(int n, const BodyProp *props, const double *X, double invh, uint32_t mask,
                           int4 *cellOf, int *bucketCnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BodyProp &b = props[i];
  if (b.large) { cellOf[i] = make_int4(0, 0, 0, -1); return; }
  const double *p = X + size_t(i) * XS;
  const int cx = int(floor(p[0] * invh)), cy = int(floor(p[1] * invh)), cz = int(floor(p[2] * invh));
  cellOf[i] = make_int4(cx, cy, cz, b.scene);
  atomicAdd(&bucketCnt[cellHash(b.scene, cx, cy, cz) & mask], 1);
}

// Bucket the small bodies; each entry carries its cell so that a bucket scan
// reads it in sequence.
__global__ void k_bp_scatter(int n, const int4 *cellOf, uint32_t mask, int *cursor, int *entries, int4 *entryCell) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int4 c = cellOf[i];
  if (c.w < 0) return;
  const int pos = atomicAdd(&cursor[cellHash(c.w, c.x, c.y, c.z) & mask], 1);
  entries[pos] = i;
  entryCell[pos] = c;
}

// Is there a force between x and y? kind: 1 live joint, 2 manifold, 0 none.
// Looks at x's force list (the CSR of the last step).
// The list is read in chunks of 8 entries (one round trip for the chunk's
// others and keys, not one per entry); matches are taken in list order.
__device__ __forceinline__ int linkKind(int x, int y, const int *adjStart, const int *adjKey, const int *adjOther,
                                        int J, const JointDev *joints) {
  constexpr int CH = 8;
  const int e1 = adjStart[x + 1];
  for (int e0 = adjStart[x]; e0 < e1; e0 += CH) {
    int o[CH], key[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
      const bool in = e0 + k < e1;
      o[k] = in ? adjOther[e0 + k] : -1;
      key[k] = in ? adjKey[e0 + k] : 0;
    }
#pragma unroll
    for (int k = 0; k < CH; k++) {
      if (o[k] != y) continue;
      if (key[k] < J) {
        if (!joints[key[k]].broken) return 1;
      } else {
        return 2;
      }
    }
  }
  return 0;
}

struct BpArgs {
  int n, J;
  const BodyProp *props;
  const double *X;
  const int4 *cellOf;
  const int *bucketStart, *bucketCnt, *entries;
  uint32_t mask;
  const int *sceneLargeStart, *sceneLarges;  // large bodies per scene
  const int *sceneStart;                     // first body of each scene
  const int *adjStart, *adjKey, *adjOther;
  const JointDev *joints;
  int *countRev, *countNewRev;               // [n] counts for A stored at n-1-A
  const int *startRev, *startNewRev;         // [n] exclusive scans of the counts
  int2 *pairs, *newPairs;                    // output lists (A, B)
  int pairCap;
  const int4 *entryCell;                     // cell of each bucket entry
  int *scratch;                              // [n * BP_CAP] per-body sorted lists
  const double *skin;                        // [n] contact offsets (GPU contact model), or null
};

// The predicate of solver.js:316-329 for A > B, plus the two flags: is the
// pair recorded (a dynamic body, no live joint), does it create a manifold
// (no manifold yet).
__device__ __forceinline__ int pairTest(const BpArgs &a, int A, int B) {
  const BodyProp &pa = a.props[A], &pb = a.props[B];
  const double *xa = a.X + size_t(A) * XS, *xb = a.X + size_t(B) * XS;
  const double dx = xa[0] - xb[0], dy = xa[1] - xb[1], dz = xa[2] - xb[2];
  double r = pa.radius + pb.radius;
  // the GPU contact model grows the spheres by the pair's skin (shaders.js:440-441)
  if (a.skin) r = r + (a.skin[A] + a.skin[B]);
  if (!(dx * dx + dy * dy + dz * dz <= r * r)) return 0;
  if (!(pa.mass > 0 || pb.mass > 0)) return 0;
  // look the link up from the side with the shorter force list
  const int la = a.adjStart[A + 1] - a.adjStart[A], lb = a.adjStart[B + 1] - a.adjStart[B];
  const int kind = la <= lb ? linkKind(A, B, a.adjStart, a.adjKey, a.adjOther, a.J, a.joints)
                            : linkKind(B, A, a.adjStart, a.adjKey, a.adjOther, a.J, a.joints);
  if (kind == 1) return 0;
  return kind == 2 ? 1 : 2;  // 1: recorded, existing manifold; 2: recorded, new
}

// pairTest with A's force list already in LDS (entries: other body in the low
// 30 bits, kind in the top two: 1 live joint, 2 manifold, 0 a broken joint,
// which links nothing). The first linking entry in list order decides, as
// linkKind does; A's list holds every force between A and B, so the answer
// is linkKind's whichever side it reads.
__device__ __forceinline__ int pairTestList(const BpArgs &a, int A, int B, const uint32_t *list, int len) {
  const BodyProp &pa = a.props[A], &pb = a.props[B];
  const double *xa = a.X + size_t(A) * XS, *xb = a.X + size_t(B) * XS;
  const double dx = xa[0] - xb[0], dy = xa[1] - xb[1], dz = xa[2] - xb[2];
  double r = pa.radius + pb.radius;
  if (a.skin) r = r + (a.skin[A] + a.skin[B]);
  if (!(dx * dx + dy * dy + dz * dz <= r * r)) return 0;
  if (!(pa.mass > 0 || pb.mass > 0)) return 0;
  uint32_t kind = 0;
  for (int k = 0; k < len; k++) {
    const uint32_t v = list[k];
    if ((v & 0x3fffffffu) == uint32_t(B) && (v >> 30) != 0u) {
      kind = v >> 30;
      break;
    }
  }
  if (kind == 1u) return 0;
  return kind == 2u ? 1 : 2;
}

// Insert (x, y) into the run list[base .. base+cnt) kept sorted by y descending.
__device__ __forceinline__ void insertDesc(int2 *list, int base, int cnt, int cap, int x, int y) {
  int k = base + cnt;
  if (k >= cap) return;
  while (k > base && list[k - 1].y < y) {
    list[k] = list[k - 1];
    k--;
  }
  list[k] = make_int2(x, y);
}

// Small bodies: pairs (x, y) with y < x from the grid, and (x, L) with the
// scene's large bodies L < x, sorted by y descending (the brute-force order
// i = N-1..1, j = i-1..0 of broadphase.js:236-243 restricted to one i).
//
// One pass with a team of 32 lanes per body: the 27 neighbour cells' bucket
// entries are numbered as one list (a team prefix sum over their counts) and
// the lanes take its entries in turn, so that a crowded cell is shared by
// the whole team rather than left to one lane; then the scene's large bodies
// likewise. The accepted pairs are
// ranked by y in LDS and written to the body's slot of a scratch list (at most
// BP_CAP entries; a body with more is listed again, serially, when the lists
// are emitted). The counts go to countRev / countNewRev (indexed n-1-x so an
// exclusive scan yields the descending-x layout).
constexpr int BP_TEAM = 32, BP_CAP = 64, BP_NEW = 0x40000000;

__global__ void __launch_bounds__(256) k_bp_team(BpArgs a) {
  constexpr int TEAMS = 256 / BP_TEAM, ADJ_LDS = 64;
  __shared__ int sy[TEAMS][BP_CAP];
  __shared__ int scount[TEAMS], snew[TEAMS];
  __shared__ uint32_t sadj[TEAMS][ADJ_LDS];
  __shared__ int sbs[TEAMS][27], spre[TEAMS][27];  // each neighbour cell's bucket start and first list index
  const int tm = threadIdx.x / BP_TEAM, lane = threadIdx.x % BP_TEAM;
  const int x = blockIdx.x * TEAMS + tm;
  if (lane == 0) { scount[tm] = 0; snew[tm] = 0; }
  // x's force list, with each entry's kind, into LDS: every pair test of
  // this body then searches it there instead of in memory
  int alen = ADJ_LDS + 1;
  if (x < a.n) {
    const int e0 = a.adjStart[x];
    alen = a.adjStart[x + 1] - e0;
    if (alen <= ADJ_LDS)
      for (int k = lane; k < alen; k += BP_TEAM) {
        const int key = a.adjKey[e0 + k];
        const uint32_t kind = key < a.J ? (a.joints[key].broken ? 0u : 1u) : 2u;
        sadj[tm][k] = (uint32_t(a.adjOther[e0 + k]) & 0x3fffffffu) | (kind << 30);
      }
  }
  const bool active = x < a.n && !a.props[x].large;
  // the neighbour cells' buckets: lane q < 27 looks up cell q, and a prefix
  // sum over the team numbers their entries as one list
  const int4 c = active ? a.cellOf[x] : make_int4(0, 0, 0, 0);
  int cntq = 0;
  if (active && lane < 27) {
    const int qx = c.x + lane % 3 - 1, qy = c.y + (lane / 3) % 3 - 1, qz = c.z + lane / 9 - 1;
    const uint32_t bkt = cellHash(c.w, qx, qy, qz) & a.mask;
    sbs[tm][lane] = a.bucketStart[bkt];
    cntq = a.bucketCnt[bkt];
  }
  int incl = cntq;
#pragma unroll
  for (int off = 1; off < BP_TEAM; off <<= 1) {
    const int u = __shfl_up(incl, off, BP_TEAM);
    if (lane >= off) incl += u;
  }
  if (lane < 27) spre[tm][lane] = incl - cntq;
  const int total = __shfl(incl, 26, BP_TEAM);
  __syncthreads();
  const bool inLds = alen <= ADJ_LDS;
  auto test = [&](int y) { return inLds ? pairTestList(a, x, y, sadj[tm], alen) : pairTest(a, x, y); };
  if (active) {
    auto accept = [&](int y, int t) {
      const int k = atomicAdd(&scount[tm], 1);
      if (t == 2) atomicAdd(&snew[tm], 1);
      if (k < BP_CAP) sy[tm][k] = y | (t == 2 ? BP_NEW : 0);
    };
    for (int j = lane; j < total; j += BP_TEAM) {
      // the cell holding list index j: the last one whose first index is at
      // most j (cells with no entries share their successor's first index)
      int q = 0;
#pragma unroll
      for (int step = 16; step > 0; step >>= 1)
        if (q + step < 27 && spre[tm][q + step] <= j) q += step;
      const int e = sbs[tm][q] + (j - spre[tm][q]);
      const int qx = c.x + q % 3 - 1, qy = c.y + (q / 3) % 3 - 1, qz = c.z + q / 9 - 1;
      const int y = a.entries[e];
      if (y >= x) continue;
      const int4 cy = a.entryCell[e];
      if (cy.x != qx || cy.y != qy || cy.z != qz || cy.w != c.w) continue;
      const int t = test(y);
      if (t) accept(y, t);
    }
    const int sc = a.props[x].scene;
    const int l0 = a.sceneLargeStart[sc], l1 = a.sceneLargeStart[sc + 1];
    for (int e = l0 + lane; e < l1; e += BP_TEAM) {
      const int y = a.sceneLarges[e];
      if (y >= x) continue;
      const int t = test(y);
      if (t) accept(y, t);
    }
  }
  __syncthreads();
  if (!active) return;
  const int cnt = scount[tm], m = min(cnt, BP_CAP);
  int *out = a.scratch + size_t(x) * BP_CAP;
  for (int j = lane; j < m; j += BP_TEAM) {
    const int v = sy[tm][j], y = v & ~BP_NEW;
    int rank = 0;
    for (int k = 0; k < m; k++) rank += (sy[tm][k] & ~BP_NEW) > y;
    out[rank] = v;
  }
  if (lane == 0) {
    a.countRev[a.n - 1 - x] = cnt;
    a.countNewRev[a.n - 1 - x] = snew[tm];
  }
}

// Serial listing of one small body straight into the output (insertion into
// the descending run): the path for a body with more than BP_CAP pairs.
__device__ void bpSerial(const BpArgs &a, int x) {
  const int4 c = a.cellOf[x];
  const int base = a.startRev[a.n - 1 - x], nbase = a.startNewRev[a.n - 1 - x];
  int cnt = 0, ncnt = 0;
  auto accept = [&](int y, int t) {
    insertDesc(a.pairs, base, cnt, a.pairCap, x, y);
    if (t == 2) insertDesc(a.newPairs, nbase, ncnt, a.pairCap, x, y);
    cnt++;
    ncnt += t == 2;
  };
  for (int dx = -1; dx <= 1; dx++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++) {
        const int qx = c.x + dx, qy = c.y + dy, qz = c.z + dz;
        const uint32_t bkt = cellHash(c.w, qx, qy, qz) & a.mask;
        const int s0 = a.bucketStart[bkt], s1 = s0 + a.bucketCnt[bkt];
        for (int e = s0; e < s1; e++) {
          const int y = a.entries[e];
          if (y >= x) continue;
          const int4 cy = a.entryCell[e];
          if (cy.x != qx || cy.y != qy || cy.z != qz || cy.w != c.w) continue;
          const int t = pairTest(a, x, y);
          if (t) accept(y, t);
        }
      }
  const int sc = a.props[x].scene;
  for (int e = a.sceneLargeStart[sc]; e < a.sceneLargeStart[sc + 1]; e++) {
    const int y = a.sceneLarges[e];
    if (y >= x) continue;
    const int t = pairTest(a, x, y);
    if (t) accept(y, t);
  }
}

// Emit each small body's sorted list at its place in the pair lists (all
// pairs, and the subsequence that opens a new manifold).
__global__ void k_bp_emit(BpArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.n || a.props[x].large) return;
  const int cnt = a.countRev[a.n - 1 - x];
  if (cnt == 0) return;
  if (cnt > BP_CAP) {
    bpSerial(a, x);
    return;
  }
  const int base = a.startRev[a.n - 1 - x], nbase = a.startNewRev[a.n - 1 - x];
  const int *in = a.scratch + size_t(x) * BP_CAP;
  int k = 0;
  for (int j = 0; j < cnt; j++) {
    const int v = in[j], y = v & ~BP_NEW;
    if (base + j < a.pairCap) a.pairs[base + j] = make_int2(x, y);
    if (v & BP_NEW) {
      if (nbase + k < a.pairCap) a.newPairs[nbase + k] = make_int2(x, y);
      k++;
    }
  }
}

// Large bodies: one 256-thread block per large body L tests every y < L of
// its scene, in descending y, and compacts the accepted ones in that order.
template <int PASS>
__global__ void k_bp_large(BpArgs a, const int *largeList) {
  const int L = largeList[blockIdx.x];
  const int lo = a.sceneStart[a.props[L].scene];
  __shared__ int wc[4], wn[4];
  __shared__ int running, runningNew;
  if (threadIdx.x == 0) running = runningNew = 0;
  __syncthreads();
  const int base = PASS == 1 ? a.startRev[a.n - 1 - L] : 0;
  const int nbase = PASS == 1 ? a.startNewRev[a.n - 1 - L] : 0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int top = L - 1; top >= lo; top -= 256) {
    const int y = top - int(threadIdx.x);
    const int t = y >= lo ? pairTest(a, L, y) : 0;
    const unsigned long long m = __ballot(t != 0), mn = __ballot(t == 2);
    const unsigned long long below = (1ull << lane) - 1ull;
    if (lane == 0) { wc[w] = __popcll(m); wn[w] = __popcll(mn); }
    __syncthreads();
    int off = running, noff = runningNew;
    for (int k = 0; k < w; k++) { off += wc[k]; noff += wn[k]; }
    if (t && PASS == 1) {
      const int pos = base + off + __popcll(m & below);
      if (pos < a.pairCap) a.pairs[pos] = make_int2(L, y);
      if (t == 2) {
        const int np = nbase + noff + __popcll(mn & below);
        if (np < a.pairCap) a.newPairs[np] = make_int2(L, y);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      running += wc[0] + wc[1] + wc[2] + wc[3];
      runningNew += wn[0] + wn[1] + wn[2] + wn[3];
    }
    __syncthreads();
  }
  if (PASS == 0 && threadIdx.x == 0) {
    a.countRev[a.n - 1 - L] = running;
    a.countNewRev[a.n - 1 - L] = runningNew;
  }
}

// ------------------------------------------------------------ narrow phase
// Manifold.initialize (manifold.js:214-318) for every old manifold (t < nOld)
// and every new pair, written to slot t of the other buffer. Under the GPU
// contact model (P.contactModel, mat and V given) it follows the reference
// GPU backend's narrowphase (shaders.js:1105-1367) instead: pair material,
// the contact skin, the geometric warm-start fallback, the rest distance and
// the restitution target.
// poly: the block's LDS clip polygons (one lane-interleaved column per thread).
template <class R, bool GL>
__device__ __forceinline__ void narrowOne(int t, int nOld, const int *liveOld, ManBuf<R> ob, ManBuf<R> nb,
                                          const int2 *newPairs, const BodyProp *props, const double *X,
                                          const StepParams &P, int *keep, int *contactTotal, int *srcSlot,
                                          const BodyMat *mat, const double *V, double *poly) {
  int a, b, m = -1, ncOld = 0;
  if (t < nOld) {
    m = liveOld[t];
    const int4 h = ob.head[m];
    a = h.x; b = h.y; ncOld = h.z;
  } else {
    const int2 pr = newPairs[t - nOld];
    a = pr.x; b = pr.y;
  }
  const BodyProp &pa = props[a], &pb = props[b];
  BodyView A{X + size_t(a) * XS, X + size_t(a) * XS + 3, pa.hx, pa.hy, pa.hz, pa.radius, pa.shape};
  BodyView B{X + size_t(b) * XS, X + size_t(b) * XS + 3, pb.hx, pb.hy, pb.hz, pb.radius, pb.shape};
  const bool hasSphere = pa.shape == 1 || pb.shape == 1;
  const bool hadContacts = ncOld > 0;
  double oldBasis[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double basis[9];
  if (m >= 0)
    for (int k = 0; k < 9; k++) basis[k] = double(ob.basis[size_t(m) * BAS + k]);
  else
    for (int k = 0; k < 9; k++) basis[k] = 0;
  constexpr bool gpu = GL;  // the GPU contact law (P.contactModel), compiled in only where launched for it
  if ((hasSphere || gpu) && hadContacts)
    for (int k = 0; k < 9; k++) oldBasis[k] = basis[k];
  // pair material: the CPU engine's one coefficient, or the GPU model's
  // static and dynamic coefficients, restitution, skin and rest distance
  double mu, muD = 0, cdist = 0, rdist = 0, rest = 0;
  if (gpu) {
    const BodyMat &ma = mat[a], &mm = mat[b];
    mu = xsqrt(ma.staticFriction * mm.staticFriction);
    muD = xsqrt(ma.dynamicFriction * mm.dynamicFriction);
    cdist = ma.contactOffset + mm.contactOffset;
    rdist = ma.restOffset + mm.restOffset;
    rest = jmax(ma.restitution, mm.restitution);
  } else {
    mu = xsqrt(pa.friction * pb.friction);
    muD = mu;
  }

  // the old contacts' feature keys, read once for every new contact's match
  // (requested before the collision test, which hides their latency)
  uint32_t ofeat[MAXC];
#pragma unroll
  for (int j = 0; j < MAXC; j++) ofeat[j] = j < ncOld ? ob.feat[size_t(m) * MAXC + j] : 0u;
  // the new contacts go straight to this manifold's output slots
  const size_t nbase = size_t(t) * MAXC;
  AliasDouble *mids = reinterpret_cast<AliasDouble *>(nb.hot + nbase * HOT);
  constexpr int MID_STRIDE = int(HOT * sizeof(R) / sizeof(double));
  const int n = collide(A, B, nb.anchor + nbase * 6, nb.feat + nbase, mids, MID_STRIDE, basis, poly + threadIdx.x,
                        cdist);
  const bool compatible =
      !hasSphere || !hadContacts || (oldBasis[0] * basis[0] + oldBasis[1] * basis[1] + oldBasis[2] * basis[2]) > 0.85;
  // GPU model: an old contact matches by feature key or geometrically, each
  // old contact once (shaders.js:1174-1319)
  const bool basisCompatible =
      hadContacts && (oldBasis[0] * basis[0] + oldBasis[1] * basis[1] + oldBasis[2] * basis[2]) > 0.85;
  double persistence = 0;
  if (gpu) {
    const double minExtent = 2.0 * jmin(jmin(pa.hx, jmin(pa.hy, pa.hz)), jmin(pb.hx, jmin(pb.hy, pb.hz)));
    persistence = jmax(P.persistence, jmax(cdist, minExtent * 0.02));
  }
  const double persistenceSq = persistence * persistence;
  uint32_t oldUsed = 0;

  // merged per-contact state: anchors, k, lambda, C, stick
  for (int i = 0; i < n; i++) {
    const double *nca = nb.anchor + (nbase + i) * 6;
    double rA[3] = {nca[0], nca[1], nca[2]};
    double rB[3] = {nca[3], nca[4], nca[5]};
    const uint32_t feature = nb.feat[nbase + i];
    double kk[3] = {0, 0, 0}, lam[3] = {0, 0, 0}, Cc[3] = {0, 0, 0};
    int stick = gpu ? 1 : 0, src = -1;  // a new GPU-model contact starts static
    bool found = false, priorOutsideRest = false, priorHadBias = false;
    int match = -1;
    if (!gpu) {
      // the first old contact with this feature key (manifold.js:241-286)
#pragma unroll
      for (int j = MAXC - 1; j >= 0; j--)
        if (j < ncOld && ofeat[j] == feature) match = j;
      if (!compatible) match = -1;
    } else if (basisCompatible) {
      // this contact's midpoint, from the merge test's record in its own
      // (not yet written) slot
      const AliasDouble *mp = mids + i * MID_STRIDE;
      const double mx = mp[0], my = mp[1], mz = mp[2];
      bool nearOne = false;
      double best = persistenceSq;
      for (int j = 0; j < ncOld; j++) {
        if (oldUsed & (1u << j)) continue;
        const bool same = ofeat[j] == feature;
        if (hasSphere && same) { match = j; break; }
        const double *oa = ob.anchor + (size_t(m) * MAXC + j) * 6;
        double oxA[3], oxB[3], od[3], tg[3];
        xform(oxA, A.p, A.q, oa);
        xform(oxB, B.p, B.q, oa + 3);
        for (int k = 0; k < 3; k++) od[k] = oxA[k] - oxB[k];
        const double sep = basis[0] * od[0] + basis[1] * od[1] + basis[2] * od[2];
        for (int k = 0; k < 3; k++) tg[k] = od[k] - basis[k] * sep;
        if (sep > cdist + persistence) continue;
        if (tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2] > persistenceSq) continue;
        const double dx = mx - (oxA[0] + oxB[0]) * 0.5, dy = my - (oxA[1] + oxB[1]) * 0.5,
                     dz = mz - (oxA[2] + oxB[2]) * 0.5;
        const double md = dx * dx + dy * dy + dz * dz;
        if (md > persistenceSq) continue;
        if (same) { match = j; break; }
        if (!nearOne || md < best) { nearOne = true; match = j; best = md; }
      }
      if (match >= 0) oldUsed |= 1u << match;
    }
    if (match >= 0) {
      found = true;
      const size_t os = size_t(m) * MAXC + match;
      {
        src = int(os);
        const R *oh = ob.hot + os * HOT;
        for (int k = 0; k < 3; k++) {
          kk[k] = double(oh[HK + k]);
          lam[k] = double(oh[HL + k]);
          Cc[k] = double(ob.Cd[os * 4 + k]);
        }
        const int oldStick = ob.stick[os];
        stick = oldStick;
        priorOutsideRest = oh[HC0] > R(0);
        priorHadBias = oh[HB] > R(0);
        if (hasSphere || gpu) {
          // transportComponents / transportPenalty (manifold.js:73-162)
          const double x = lam[0], y = lam[1], z = lam[2];
          const double px = kk[0], py = kk[1], pz = kk[2];
          const double *o = oldBasis, *q = basis;
          for (int r = 0; r < 3; r++) {
            const int rr = r * 3;
            lam[r] = x * (q[rr] * o[0] + q[rr + 1] * o[1] + q[rr + 2] * o[2]) +
                     y * (q[rr] * o[3] + q[rr + 1] * o[4] + q[rr + 2] * o[5]) +
                     z * (q[rr] * o[6] + q[rr + 1] * o[7] + q[rr + 2] * o[8]);
          }
          for (int r = 0; r < 3; r++) {
            const int rr = r * 3;
            const double ddx = q[rr] * o[0] + q[rr + 1] * o[1] + q[rr + 2] * o[2];
            const double ddy = q[rr] * o[3] + q[rr + 1] * o[4] + q[rr + 2] * o[5];
            const double ddz = q[rr] * o[6] + q[rr + 1] * o[7] + q[rr + 2] * o[8];
            kk[r] = px * ddx * ddx + py * ddy * ddy + pz * ddz * ddz;
          }
        }
        if (oldStick && !hasSphere) {
          const double *an = ob.anchor + os * 6;
          for (int k = 0; k < 3; k++) { rA[k] = an[k]; rB[k] = an[3 + k]; }
          src |= 0x40000000;  // anchors kept
        }
      }
    }
    // C0 at x_t and the warm start (manifold.js:290-315)
    double xa[3], xb[3], d[3], C0[3];
    xform(xa, A.p, A.q, rA);
    xform(xb, B.p, B.q, rB);
    d[0] = xa[0] - xb[0]; d[1] = xa[1] - xb[1]; d[2] = xa[2] - xb[2];
    mulv(C0, basis, d);
    double bias = 0;
    if (gpu) {
      // C0 against the rest distance; the restitution target of a new impact
      // (or of one a speculative-only contact kept armed) whose closing
      // motion crosses the rest surface this step (shaders.js:1322-1350)
      C0[0] = C0[0] - rdist;
      if ((!found || (priorOutsideRest && !priorHadBias)) && rest > 0.0) {
        const double *va = V + size_t(a) * VS, *vb = V + size_t(b) * VS;
        double ra[3], rb[3], ca[3], cb[3], rel[3];
        qrot(ra, A.q, rA);
        qrot(rb, B.q, rB);
        cross3(ca, va + 3, ra);
        cross3(cb, vb + 3, rb);
        for (int k = 0; k < 3; k++) rel[k] = (va[k] + ca[k]) - (vb[k] + cb[k]);
        const double vn = basis[0] * rel[0] + basis[1] * rel[1] + basis[2] * rel[2];
        const double closing = jmax(0.0, -vn * P.dt);
        const double toRest = jmax(0.0, C0[0]);
        if (vn < -P.restThreshold && closing > toRest) bias = rest * (closing - toRest);
      }
    } else {
      C0[0] += P.margin;
    }
    if (!P.postStab) {
      for (int k = 0; k < 3; k++) lam[k] = lam[k] * P.alpha;
      for (int k = 0; k < 3; k++) lam[k] = lam[k] * P.gamma;
    }
    for (int k = 0; k < 3; k++) kk[k] = kk[k] * P.gamma;
    for (int k = 0; k < 3; k++) kk[k] = clampj(kk[k], P.pmin, P.pmax);
    // cached angular Jacobian anchors at x_t (manifold.js:338-349)
    double rAw[3], rBw[3];
    qrot(rAw, A.q, rA);
    qrot(rBw, B.q, rB);
    const size_t s = nbase + i;
    R *h = nb.hot + s * HOT;
    for (int k = 0; k < 3; k++) {
      h[HC0 + k] = R(C0[k]); h[HK + k] = R(kk[k]); h[HL + k] = R(lam[k]);
      h[HRA + k] = R(rAw[k]); h[HRB + k] = R(rBw[k]);
    }
    h[HB] = R(bias);
    double *an = nb.anchor + s * 6;
    for (int k = 0; k < 3; k++) { an[k] = rA[k]; an[3 + k] = rB[k]; }
    nb.stick[s] = stick;
    srcSlot[s] = src;
    for (int k = 0; k < 3; k++) nb.Cd[s * 4 + k] = R(Cc[k]);
    nb.Cd[s * 4 + 3] = R(0);
  }
  nb.head[t] = make_int4(a, b, n, hasSphere ? int(MF_SPHERE) : 0);
  for (int k = 0; k < 9; k++) nb.basis[size_t(t) * BAS + k] = R(basis[k]);
  nb.basis[size_t(t) * BAS + BMU] = R(mu);
  nb.basis[size_t(t) * BAS + BMUD] = R(gpu ? muD : 0.0);
  nb.basis[size_t(t) * BAS + BCD] = R(cdist);
  keep[t] = n > 0 ? 1 : 0;
  if (n > 0) atomicAdd(contactTotal, n);
}

// One thread per manifold, 64 threads a block. counts (device, or null): the
// old and new counts as a step on the device produced them, for a launch
// whose grid was sized before they were known; the threads then stride.
template <class R, bool GL>
__global__ void __launch_bounds__(64) k_narrow(int nOld, int nNew, const int *counts, const int *liveOld, ManBuf<R> ob,
                                               ManBuf<R> nb, const int2 *newPairs, const BodyProp *props,
                                               const double *X, StepParams P, int *keep, int *contactTotal,
                                               int *srcSlot, const BodyMat *mat, const double *V) {
  __shared__ double poly[2 * POLY_LDS * 3 * LDS_LANES];
  if (counts) { nOld = counts[0]; nNew = counts[1]; }
  const int nTot = nOld + nNew;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nTot; t += gridDim.x * blockDim.x)
    narrowOne<R, GL>(t, nOld, liveOld, ob, nb, newPairs, props, X, P, keep, contactTotal, srcSlot, mat, V, poly);
}

// ----------------------------------------------------------------- joints
__device__ const double IDQ[4] = {0, 0, 0, 1};

// joint.js:115-128
__device__ __forceinline__ void jointLin(const JointDev &j, const double *X, double *o) {
  double pa[3], pb[3];
  if (j.a >= 0) xform(pa, X + size_t(j.a) * XS, X + size_t(j.a) * XS + 3, j.rA);
  else { pa[0] = j.rA[0]; pa[1] = j.rA[1]; pa[2] = j.rA[2]; }
  xform(pb, X + size_t(j.b) * XS, X + size_t(j.b) * XS + 3, j.rB);
  o[0] = pa[0] - pb[0]; o[1] = pa[1] - pb[1]; o[2] = pa[2] - pb[2];
}
__device__ __forceinline__ void jointAng(const JointDev &j, const double *X, double *o) {
  qsub(o, j.a >= 0 ? X + size_t(j.a) * XS + 3 : IDQ, X + size_t(j.b) * XS + 3);
  o[0] = o[0] * j.torqueArm; o[1] = o[1] * j.torqueArm; o[2] = o[2] * j.torqueArm;
}

// spring.js:107-118 (_evaluate): d = pA - pB, n = d / |d|, C = |d| - rest;
// false when the anchors (nearly) coincide.
__device__ __forceinline__ bool springEval(const JointDev &j, const double *X, double *d, double *n, double &C) {
  double pa[3], pb[3];
  xform(pa, X + size_t(j.a) * XS, X + size_t(j.a) * XS + 3, j.rA);
  xform(pb, X + size_t(j.b) * XS, X + size_t(j.b) * XS + 3, j.rB);
  d[0] = pa[0] - pb[0]; d[1] = pa[1] - pb[1]; d[2] = pa[2] - pb[2];
  const double len = xsqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (len <= 1.0e-6) return false;
  n[0] = d[0] / len; n[1] = d[1] / len; n[2] = d[2] / len;
  C = len - j.stiffAng;
  return true;
}
// spring.js:121-131 (_updateTearing): true when it tears now
__device__ __forceinline__ bool springTear(JointDev &j, double C) {
  if (j.stiffAng <= 1.0e-6) return false;
  const double strain = jmax(0.0, C / j.stiffAng);
  j.lamLin[0] = jmax(j.lamLin[0], strain);
  if (strain <= j.fracture) return false;
  j.broken = 1;
  j.kLin[0] = 0;
  j.CLin[0] = 0;
  return true;
}

// joint.js:131-160; a broken joint leaves the force list. A spring
// (spring.js:133-166, stiffness ramp off as solver.js:187 ships it) takes its
// material stiffness as the penalty; a torn one stays, inactive, and one of
// stiffness 0 leaves.
__global__ void k_joint_init(int J, JointDev *joints, const double *X, StepParams P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= J) return;
  JointDev &j = joints[i];
  if (j.kind == AVBD_JOINT_SPRING) {
    if (j.broken == 2) return;
    if (j.broken) { j.kLin[0] = 0; j.CLin[0] = 0; return; }
    if (__builtin_isfinite(j.fracture)) {
      double d[3], n[3], C;
      if (springEval(j, X, d, n, C)) {
        j.CLin[0] = C;
        if (springTear(j, C)) return;
      }
    }
    j.kLin[0] = j.stiffLin;
    if (!(j.stiffLin > 0)) j.broken = 2;
    return;
  }
  if (j.broken) { j.broken = 2; return; }  // 2: removed
  jointLin(j, X, j.C0Lin);
  jointAng(j, X, j.C0Ang);
  if (!P.postStab) {
    for (int k = 0; k < 3; k++) { j.lamLin[k] = j.lamLin[k] * P.alpha; }
    for (int k = 0; k < 3; k++) { j.lamLin[k] = j.lamLin[k] * P.gamma; }
    for (int k = 0; k < 3; k++) { j.lamAng[k] = j.lamAng[k] * P.alpha; }
    for (int k = 0; k < 3; k++) { j.lamAng[k] = j.lamAng[k] * P.gamma; }
  }
  for (int k = 0; k < 3; k++) j.kLin[k] = j.kLin[k] * P.gamma;
  for (int k = 0; k < 3; k++) j.kLin[k] = clampj(j.kLin[k], P.pmin, P.pmax);
  for (int k = 0; k < 3; k++) j.kAng[k] = j.kAng[k] * P.gamma;
  for (int k = 0; k < 3; k++) j.kAng[k] = clampj(j.kAng[k], P.pmin, P.pmax);
  for (int k = 0; k < 3; k++) j.kLin[k] = jmin(j.kLin[k], j.stiffLin);
  for (int k = 0; k < 3; k++) j.kAng[k] = jmin(j.kAng[k], j.stiffAng);
}

// Full-matrix helpers for the joint rows (their products are written out as
// the reference writes them).
__device__ __forceinline__ void mT(double *o, const double *a) {
  o[0] = a[0]; o[1] = a[3]; o[2] = a[6]; o[3] = a[1]; o[4] = a[4]; o[5] = a[7]; o[6] = a[2]; o[7] = a[5]; o[8] = a[8];
}
__device__ __forceinline__ void mDiagR(double *o, const double *a, const double *d) {
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) o[r * 3 + c] = a[r * 3 + c] * d[c];
}
__device__ __forceinline__ void addLower(double *lt, const double *m) {
  lt[0] += m[0]; lt[1] += m[3]; lt[2] += m[4]; lt[3] += m[6]; lt[4] += m[7]; lt[5] += m[8];
}

__device__ __forceinline__ void outer3(double *o, const double *a, const double *b) {
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) o[r * 3 + c] = a[r] * b[c];
}

// Spring.updatePrimal (spring.js:168-212) with the geometric stiffness of
// Equation 17 lumped by full 6-vector column norms (spring.js:214-269, on as
// shipped: solver.js:190-193).
__device__ void springStamp(Sys<double> &s, const JointDev &j, bool isA, const double *X) {
  if (j.broken) return;
  double d[3], n[3], C;
  if (!springEval(j, X, d, n, C)) return;
  const double k = j.kLin[0];
  const double fo = k * C;
  double rW[3], jl[3], ja[3], m0[9];
  if (isA) {
    qrot(rW, X + size_t(j.a) * XS + 3, j.rA);
    jl[0] = n[0]; jl[1] = n[1]; jl[2] = n[2];
    cross3(ja, rW, n);
  } else {
    qrot(rW, X + size_t(j.b) * XS + 3, j.rB);
    jl[0] = -n[0]; jl[1] = -n[1]; jl[2] = -n[2];
    cross3(ja, rW, n);
    ja[0] = -ja[0]; ja[1] = -ja[1]; ja[2] = -ja[2];
  }
  outer3(m0, jl, jl);
  for (int e = 0; e < 9; e++) m0[e] = m0[e] * k;
  addLower(s.lin, m0);
  outer3(m0, ja, ja);
  for (int e = 0; e < 9; e++) m0[e] = m0[e] * k;
  addLower(s.ang, m0);
  outer3(m0, ja, jl);
  for (int e = 0; e < 9; e++) s.cross[e] += m0[e] * k;
  for (int e = 0; e < 3; e++) s.rl[e] += jl[e] * fo;
  for (int e = 0; e < 3; e++) s.ra[e] += ja[e] * fo;

  const double len = xsqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double sg = isA ? 1.0 : -1.0;
  double Pm[9], Sk[9], Gla[9], Gal[9], Gaa[9];
  outer3(Pm, n, n);
  for (int e = 0; e < 9; e++) Pm[e] = -Pm[e];
  Pm[0] += 1; Pm[4] += 1; Pm[8] += 1;
  for (int e = 0; e < 9; e++) Pm[e] = Pm[e] / len;
  Sk[0] = 0; Sk[1] = -rW[2]; Sk[2] = rW[1]; Sk[3] = rW[2]; Sk[4] = 0; Sk[5] = -rW[0]; Sk[6] = -rW[1]; Sk[7] = rW[0];
  Sk[8] = 0;
  mul33(Gla, Pm, Sk);
  for (int e = 0; e < 9; e++) Gla[e] = -Gla[e];
  mul33(Gal, Sk, Pm);
  mul33(Gaa, Sk, Pm);
  mul33(Gaa, Gaa, Sk);
  for (int e = 0; e < 9; e++) Gaa[e] = -Gaa[e];
  const double ndr = n[0] * rW[0] + n[1] * rW[1] + n[2] * rW[2];
  outer3(m0, rW, n);
  m0[0] -= ndr; m0[4] -= ndr; m0[8] -= ndr;
  for (int e = 0; e < 9; e++) Gaa[e] += m0[e] * sg;
  const double af = xabs(fo);
  const int diag[3] = {0, 2, 5};
  for (int c = 0; c < 3; c++) {
    const double a0 = Pm[c], a1 = Pm[3 + c], a2 = Pm[6 + c], a3 = Gal[c], a4 = Gal[3 + c], a5 = Gal[6 + c];
    s.lin[diag[c]] += xsqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4 + a5 * a5) * af;
  }
  for (int c = 0; c < 3; c++) {
    const double a0 = Gla[c], a1 = Gla[3 + c], a2 = Gla[6 + c], a3 = Gaa[c], a4 = Gaa[3 + c], a5 = Gaa[6 + c];
    s.ang[diag[c]] += xsqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4 + a5 * a5) * af;
  }
}

// Joint.updatePrimal (joint.js:162-253) into the body's system.
__device__ void jointStamp(Sys<double> &s, const JointDev &j, bool isA, const double *X, double alpha) {
  if (j.kind == AVBD_JOINT_SPRING) {
    springStamp(s, j, isA, X);
    return;
  }
  const double INF = __builtin_inf();
  if (j.kLin[0] * j.kLin[0] + j.kLin[1] * j.kLin[1] + j.kLin[2] * j.kLin[2] > 0) {
    double K[9] = {j.kLin[0], 0, 0, 0, j.kLin[1], 0, 0, 0, j.kLin[2]};
    double C[3], F[3], jl[9], ja[9], r[3], rW[3];
    jointLin(j, X, C);
    if (j.stiffLin == INF)
      for (int k = 0; k < 3; k++) C[k] += j.C0Lin[k] * -alpha;
    mulv(F, K, C);
    for (int k = 0; k < 3; k++) F[k] = F[k] + j.lamLin[k];
    if (isA) {
      for (int k = 0; k < 9; k++) jl[k] = (k % 4 == 0) ? 1.0 : 0.0;
      qrot(rW, X + size_t(j.a) * XS + 3, j.rA);
      r[0] = -rW[0]; r[1] = -rW[1]; r[2] = -rW[2];
      ja[0] = 0; ja[1] = -r[2]; ja[2] = r[1]; ja[3] = r[2]; ja[4] = 0; ja[5] = -r[0]; ja[6] = -r[1]; ja[7] = r[0]; ja[8] = 0;
      r[0] = rW[0]; r[1] = rW[1]; r[2] = rW[2];
    } else {
      for (int k = 0; k < 9; k++) jl[k] = (k % 4 == 0) ? -1.0 : 0.0;
      qrot(rW, X + size_t(j.b) * XS + 3, j.rB);
      ja[0] = 0; ja[1] = -rW[2]; ja[2] = rW[1]; ja[3] = rW[2]; ja[4] = 0; ja[5] = -rW[0]; ja[6] = -rW[1]; ja[7] = rW[0]; ja[8] = 0;
      r[0] = -rW[0]; r[1] = -rW[1]; r[2] = -rW[2];
    }
    double jlT[9], jaT[9], jaTk[9], m0[9], acc[3];
    mT(jlT, jl);
    mT(jaT, ja);
    mDiagR(jaTk, jaT, j.kLin);
    mDiagR(m0, jlT, j.kLin);
    mul33(m0, m0, jl);
    addLower(s.lin, m0);
    mul33(m0, jaTk, ja);
    addLower(s.ang, m0);
    mul33(m0, jaTk, jl);
    for (int k = 0; k < 9; k++) s.cross[k] += m0[k];
    // geometric stiffness, lumped by column norms (joint.js:224-234, maths.js:373-381)
    double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, G[9];
    for (int k = 0; k < 3; k++) {
      for (int e = 0; e < 9; e++) G[e] = 0;
      G[0] = -r[k]; G[4] = -r[k]; G[8] = -r[k];
      G[0 * 3 + k] += r[0];
      G[1 * 3 + k] += r[1];
      G[2 * 3 + k] += r[2];
      for (int e = 0; e < 9; e++) G[e] = G[e] * F[k];
      for (int e = 0; e < 9; e++) H[e] += G[e];
    }
    s.ang[0] += xsqrt(H[0] * H[0] + H[3] * H[3] + H[6] * H[6]);
    s.ang[2] += xsqrt(H[1] * H[1] + H[4] * H[4] + H[7] * H[7]);
    s.ang[5] += xsqrt(H[2] * H[2] + H[5] * H[5] + H[8] * H[8]);
    mulv(acc, jlT, F);
    for (int k = 0; k < 3; k++) s.rl[k] = s.rl[k] + acc[k];
    mulv(acc, jaT, F);
    for (int k = 0; k < 3; k++) s.ra[k] = s.ra[k] + acc[k];
  }
  if (j.kAng[0] * j.kAng[0] + j.kAng[1] * j.kAng[1] + j.kAng[2] * j.kAng[2] > 0) {
    double K[9] = {j.kAng[0], 0, 0, 0, j.kAng[1], 0, 0, 0, j.kAng[2]};
    double C[3], F[3], ja[9], jaT[9], m0[9], acc[3];
    jointAng(j, X, C);
    if (j.stiffAng == INF)
      for (int k = 0; k < 3; k++) C[k] += j.C0Ang[k] * -alpha;
    mulv(F, K, C);
    for (int k = 0; k < 3; k++) F[k] = F[k] + j.lamAng[k];
    const double sg = isA ? j.torqueArm : -j.torqueArm;
    for (int k = 0; k < 9; k++) ja[k] = (k % 4 == 0) ? sg : 0.0;
    mT(jaT, ja);
    mul33(m0, jaT, K);
    mul33(m0, m0, ja);
    addLower(s.ang, m0);
    mulv(acc, jaT, F);
    for (int k = 0; k < 3; k++) s.ra[k] = s.ra[k] + acc[k];
  }
}

// springStamp for the throughput precisions: the same sums with the outer
// products written out, one reciprocal of the length, P = (I - n n^T) / len
// symmetric so that Gla = -(P S) = (S P)^T = Gal^T, Gaa = -(Gal S) plus the
// anchor term, and multiply-adds contracted.
__device__ void springStampF(Sys<double> &s, const JointDev &j, bool isA, const double *X) {
#pragma clang fp contract(fast)
  if (j.broken) return;
  double d[3], n[3], C;
  if (!springEval(j, X, d, n, C)) return;
  const double k = j.kLin[0], fo = k * C;
  double w[3], wn[3];
  if (isA) qrot(w, X + size_t(j.a) * XS + 3, j.rA);
  else qrot(w, X + size_t(j.b) * XS + 3, j.rB);
  cross3(wn, w, n);
  const double sg = isA ? 1.0 : -1.0;
  double jl[3], ja[3];
  for (int e = 0; e < 3; e++) { jl[e] = sg * n[e]; ja[e] = sg * wn[e]; }
  for (int q = 0; q < 6; q++) {
    const int r = LT_R[q], c = LT_C[q];
    s.lin[q] += jl[r] * jl[c] * k;
    s.ang[q] += ja[r] * ja[c] * k;
  }
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) s.cross[r * 3 + c] += ja[r] * jl[c] * k;
  for (int e = 0; e < 3; e++) { s.rl[e] += jl[e] * fo; s.ra[e] += ja[e] * fo; }
  // geometric stiffness of Equation 17, lumped by column norms
  const double il = 1.0 / xsqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double P[9], Sk[9], Gal[9], Gaa[9];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) P[r * 3 + c] = ((r == c ? 1.0 : 0.0) - n[r] * n[c]) * il;
  Sk[0] = 0; Sk[1] = -w[2]; Sk[2] = w[1]; Sk[3] = w[2]; Sk[4] = 0; Sk[5] = -w[0]; Sk[6] = -w[1]; Sk[7] = w[0]; Sk[8] = 0;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++)
      Gal[r * 3 + c] = Sk[r * 3 + 0] * P[c] + Sk[r * 3 + 1] * P[3 + c] + Sk[r * 3 + 2] * P[6 + c];
  const double ndw = n[0] * w[0] + n[1] * w[1] + n[2] * w[2];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++)
      Gaa[r * 3 + c] = -(Gal[r * 3 + 0] * Sk[c] + Gal[r * 3 + 1] * Sk[3 + c] + Gal[r * 3 + 2] * Sk[6 + c]) +
                       sg * (w[r] * n[c] - (r == c ? ndw : 0.0));
  const double af = xabs(fo);
  const int diag[3] = {0, 2, 5};
  for (int c = 0; c < 3; c++) {
    double lc = 0, ac = 0;
    for (int r = 0; r < 3; r++) {
      lc += P[r * 3 + c] * P[r * 3 + c] + Gal[r * 3 + c] * Gal[r * 3 + c];
      ac += Gal[c * 3 + r] * Gal[c * 3 + r] + Gaa[r * 3 + c] * Gaa[r * 3 + c];
    }
    s.lin[diag[c]] += xsqrt(lc) * af;
    s.ang[diag[c]] += xsqrt(ac) * af;
  }
}

// jointStamp for the throughput precisions: the same sums with the
// identity and diagonal factors of joint.js:162-253 multiplied out and
// multiply-adds contracted (the exact mode keeps jointStamp's order). With
// S = [rW]x, the anchor's cross-product matrix, and sigma = -1 for body A,
// +1 for body B: J_lin = -sigma I, J_ang = sigma S, so
//   Lin += K, Ang += S^T K S, Cross += J_ang^T K J_lin = -S^T K,
//   rl += -sigma F, ra += -sigma (rW x F),
// and the geometric term H = -(r.F) I + r F^T with r = -sigma rW (lumped by
// column norms into Ang's diagonal). The angular row: J = s I (s = +-the
// torque arm), Ang += s^2 K, ra += s F. Springs take springStampF.
__device__ void jointStampF(Sys<double> &s, const JointDev &j, bool isA, const double *X, double alpha) {
#pragma clang fp contract(fast)
  if (j.kind == AVBD_JOINT_SPRING) {
    springStampF(s, j, isA, X);
    return;
  }
  const double INF = __builtin_inf();
  const double *kl = j.kLin;
  if (kl[0] * kl[0] + kl[1] * kl[1] + kl[2] * kl[2] > 0) {
    double C[3], F[3], w[3];
    jointLin(j, X, C);
    if (j.stiffLin == INF)
      for (int k = 0; k < 3; k++) C[k] += j.C0Lin[k] * -alpha;
    for (int k = 0; k < 3; k++) F[k] = kl[k] * C[k] + j.lamLin[k];
    if (isA) qrot(w, X + size_t(j.a) * XS + 3, j.rA);
    else qrot(w, X + size_t(j.b) * XS + 3, j.rB);
    const double sig = isA ? -1.0 : 1.0;
    s.lin[0] += kl[0]; s.lin[2] += kl[1]; s.lin[5] += kl[2];
    s.ang[0] += kl[1] * w[2] * w[2] + kl[2] * w[1] * w[1];
    s.ang[1] += -(kl[2] * w[0] * w[1]);
    s.ang[2] += kl[0] * w[2] * w[2] + kl[2] * w[0] * w[0];
    s.ang[3] += -(kl[1] * w[0] * w[2]);
    s.ang[4] += -(kl[0] * w[1] * w[2]);
    s.ang[5] += kl[0] * w[1] * w[1] + kl[1] * w[0] * w[0];
    s.cross[1] += -(w[2] * kl[1]); s.cross[2] += w[1] * kl[2];
    s.cross[3] += w[2] * kl[0];    s.cross[5] += -(w[0] * kl[2]);
    s.cross[6] += -(w[1] * kl[0]); s.cross[7] += w[0] * kl[1];
    double rxF[3];
    cross3(rxF, w, F);
    for (int k = 0; k < 3; k++) {
      s.rl[k] += -sig * F[k];
      s.ra[k] += -sig * rxF[k];
    }
    // geometric stiffness (joint.js:224-234): columns of H = -(r.F) I + r F^T
    const double r0 = -sig * w[0], r1 = -sig * w[1], r2 = -sig * w[2];
    const double rF = r0 * F[0] + r1 * F[1] + r2 * F[2];
    const int diag[3] = {0, 2, 5};
    for (int c = 0; c < 3; c++) {
      const double h0 = r0 * F[c] - (c == 0 ? rF : 0.0), h1 = r1 * F[c] - (c == 1 ? rF : 0.0),
                   h2 = r2 * F[c] - (c == 2 ? rF : 0.0);
      s.ang[diag[c]] += xsqrt(h0 * h0 + h1 * h1 + h2 * h2);
    }
  }
  const double *ka = j.kAng;
  if (ka[0] * ka[0] + ka[1] * ka[1] + ka[2] * ka[2] > 0) {
    double C[3];
    jointAng(j, X, C);
    if (j.stiffAng == INF)
      for (int k = 0; k < 3; k++) C[k] += j.C0Ang[k] * -alpha;
    const double sg = isA ? j.torqueArm : -j.torqueArm, sg2 = sg * sg;
    s.ang[0] += sg2 * ka[0]; s.ang[2] += sg2 * ka[1]; s.ang[5] += sg2 * ka[2];
    for (int k = 0; k < 3; k++) s.ra[k] += sg * (ka[k] * C[k] + j.lamAng[k]);
  }
}

// Joint.updateDual (joint.js:255-317)
__device__ __forceinline__ void jointDualOne(JointDev &j, const double *X, const StepParams &P, double alpha) {
  if (j.broken == 2) return;
  if (j.kind == AVBD_JOINT_SPRING) {
    // Spring.updateDual (spring.js:271-288): the constraint value and
    // tearing; with the ramp off there is no penalty update
    if (j.broken) return;
    double d[3], n[3], C;
    if (!springEval(j, X, d, n, C)) return;
    j.CLin[0] = C;
    if (__builtin_isfinite(j.fracture)) springTear(j, C);
    return;
  }
  const double INF = __builtin_inf();
  if (j.kLin[0] * j.kLin[0] + j.kLin[1] * j.kLin[1] + j.kLin[2] * j.kLin[2] > 0) {
    double K[9] = {j.kLin[0], 0, 0, 0, j.kLin[1], 0, 0, 0, j.kLin[2]}, C[3], F[3];
    jointLin(j, X, C);
    if (j.stiffLin == INF) {
      for (int k = 0; k < 3; k++) C[k] += j.C0Lin[k] * -alpha;
      mulv(F, K, C);
      for (int k = 0; k < 3; k++) j.lamLin[k] = F[k] + j.lamLin[k];
    }
    for (int k = 0; k < 3; k++) j.CLin[k] = C[k];
    const double cap = jmin(j.stiffLin, P.pmax);
    for (int k = 0; k < 3; k++) j.kLin[k] = jmin(j.kLin[k] + xabs(C[k]) * P.betaLin, cap);
  }
  if (j.kAng[0] * j.kAng[0] + j.kAng[1] * j.kAng[1] + j.kAng[2] * j.kAng[2] > 0) {
    double K[9] = {j.kAng[0], 0, 0, 0, j.kAng[1], 0, 0, 0, j.kAng[2]}, C[3], F[3];
    jointAng(j, X, C);
    if (j.stiffAng == INF) {
      for (int k = 0; k < 3; k++) C[k] += j.C0Ang[k] * -alpha;
      mulv(F, K, C);
      for (int k = 0; k < 3; k++) j.lamAng[k] = F[k] + j.lamAng[k];
    }
    for (int k = 0; k < 3; k++) j.CAng[k] = C[k];
    const double cap = jmin(j.stiffAng, P.pmax);
    for (int k = 0; k < 3; k++) j.kAng[k] = jmin(j.kAng[k] + xabs(C[k]) * P.betaAng, cap);
  }
  const double l2 = j.lamAng[0] * j.lamAng[0] + j.lamAng[1] * j.lamAng[1] + j.lamAng[2] * j.lamAng[2];
  if (l2 > j.fracture * j.fracture) {
    for (int k = 0; k < 3; k++) j.kLin[k] = j.kAng[k] = j.lamLin[k] = j.lamAng[k] = 0.0;
    j.broken = 1;
  }
}

__global__ void k_joint_dual(int J, JointDev *joints, const double *X, StepParams P, double alpha) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < J) jointDualOne(joints[i], X, P, alpha);
}

// ---------------------------------------------------------------- topology
// nTotDev (or null): the count as produced on the device; the threads stride.
__global__ void k_live_compact(int nTot, const int *nTotDev, const int *keep, const int *pos, int *live) {
  if (nTotDev) nTot = *nTotDev;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nTot; t += gridDim.x * blockDim.x)
    if (keep[t]) live[pos[t]] = t;
}

// Per-body force lists. Key = j for joint j, J + k for live manifold k, so
// ascending keys are the force list's creation order. nLiveDev (or null): the
// live count as produced on the device; the threads stride.
// ctr[idx] += 1 for every lane with inc, returning each lane's old value
// (its slot). Lanes with agg share one atomic per distinct idx in the wave
// (a static body, the ground or a wall, meets hundreds of forces, and an
// atomic per force on its one counter serialises); the others add alone.
// Every lane of the wave calls it.
__device__ __forceinline__ int waveSlot(int *ctr, int idx, bool inc, bool agg) {
  int r = 0;
  if (inc && !agg) r = atomicAdd(&ctr[idx], 1);
  unsigned long long pending = __ballot(inc && agg);
  const int lane = __lane_id();
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int lidx = __shfl(idx, leader);
    const bool mine = inc && agg && idx == lidx;
    const unsigned long long same = __ballot(mine);
    int base = 0;
    if (lane == leader) base = atomicAdd(&ctr[lidx], __popcll(same));
    base = __shfl(base, leader);
    if (mine) r = base + __popcll(same & ((1ull << lane) - 1ull));
    pending &= ~same;
  }
  return r;
}

// Force lists: PASS 0 counts each body's forces (deg), PASS 1 places them
// (cursor from the scan of deg): one thread per force (joints first, then the
// live manifolds), its key and other body at a slot of each of its bodies'
// lists. The threads stride in whole-grid rounds, so a wave's lanes take
// part in every round together.
template <int PASS>
__global__ void k_adj(int nLive, const int *nLiveDev, int J, const int *live, const int4 *head,
                      const JointDev *joints, const BodyProp *props, int *deg, int *cursor, int *adjKey,
                      int *adjOther, int *owner) {
  if (nLiveDev) nLive = *nLiveDev;
  const int total = J + nLive, stride = gridDim.x * blockDim.x;
  for (int t0 = blockIdx.x * blockDim.x; t0 < total; t0 += stride) {
    const int t = t0 + int(threadIdx.x);
    int a = -1, b = -1;
    bool valid = t < total;
    if (valid) {
      if (t < J) {
        if (joints[t].broken == 2) valid = false;
        a = joints[t].a; b = joints[t].b;
      } else {
        const int4 h = head[live[t - J]];
        a = h.x; b = h.y;
      }
    }
    const bool useA = valid && a >= 0, useB = valid;
    const bool statA = useA && !(props[a].mass > 0), statB = useB && !(props[b].mass > 0);
    const int pa = waveSlot(PASS == 0 ? deg : cursor, a, useA, statA);
    const int pb = waveSlot(PASS == 0 ? deg : cursor, b, useB, statB);
    if (PASS == 1) {
      if (useA) { adjKey[pa] = t; adjOther[pa] = b; owner[pa] = a; }
      if (useB) { adjKey[pb] = t; adjOther[pb] = a; owner[pb] = b; }
    }
  }
}

// Each dynamic body's list in key order, one thread per entry: an entry's
// place is the number of smaller keys in its list (keys are distinct within a
// list), read eight at a time. Static bodies' lists keep their fill order
// (the ground's can hold thousands of entries).
__device__ __forceinline__ void adjRankOne(int p, const BodyProp *props, const int *adjStart, const int *owner,
                                           const int *keyIn, const int *otherIn, int *keyOut, int *otherOut) {
  const int b = owner[p];
  const int s0 = adjStart[b], s1 = adjStart[b + 1];
  const int key = keyIn[p];
  int rank = p - s0;
  if (props[b].mass > 0) {
    rank = 0;
    constexpr int CH = 8;
    for (int f0 = s0; f0 < s1; f0 += CH) {
      int kk[CH];
#pragma unroll
      for (int k = 0; k < CH; k++) kk[k] = f0 + k < s1 ? keyIn[f0 + k] : key;
#pragma unroll
      for (int k = 0; k < CH; k++) rank += kk[k] < key;
    }
  }
  keyOut[s0 + rank] = key;
  otherOut[s0 + rank] = otherIn[p];
}

// Entries up to min(maxEntries, the list total); the threads stride.
__global__ void k_adj_rank(int maxEntries, int n, const BodyProp *props, const int *adjStart, const int *owner,
                           const int *keyIn, const int *otherIn, int *keyOut, int *otherOut) {
  const int total = min(maxEntries, adjStart[n]);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x)
    adjRankOne(p, props, adjStart, owner, keyIn, otherIn, keyOut, otherOut);
}

// Greedy colouring (coloring.js:75-92): colour(i) is the smallest colour not
// used by a dynamic neighbour that claims before i. Claim order: body index
// (order 0, the CPU engine), or lowbias32 priority with ties to the lower
// index (order 1, layout.js:50-57 and shaders.js:1571-1585). The greedy
// result is the unique fixed point of colour(i) = mex(colours of earlier
// neighbours); relaxation reaches it from any start, including last step's.
constexpr int MAX_COLORS = 256;
// the colouring's counters (cmeta): [0..2] change counters, [3] rounds, [4]
// colours, [8..8+MAX_COLORS] colour starts, then each colour's heavy-body
// count and the heavy threshold (k_color_starts)
constexpr int CMETA_HEAVY = 8 + MAX_COLORS + 1, CMETA_N = CMETA_HEAVY + MAX_COLORS + 1;

__device__ __forceinline__ uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// a and b belong to one scene whose first body is `base`; the priority hashes
// the index within the scene, so a scene colours the same alone or batched.
__device__ __forceinline__ bool claimsBefore(int a, int b, int order, int base, const int *rank) {
  if (order == 0) return a < b;
  if (order == 2) return rank[a] < rank[b] || (rank[a] == rank[b] && a < b);
  const uint32_t pa = lowbias32(uint32_t(a - base)), pb = lowbias32(uint32_t(b - base));
  return pa < pb || (pa == pb && a < b);
}

// One relaxation round, in place: each body takes the mex of its earlier
// neighbours' current colours. Rounds repeat until one changes nothing; that
// round saw only final colours, so every body equals the mex of its earlier
// neighbours: the fixed point. ctl[0] counts this round's changes, ctl[1] is
// set once a round changed nothing (later rounds return at once), ctl[3]
// receives the number of rounds.
__global__ void k_color_round(int n, const BodyProp *props, const int *adjStart, const int *adjOther, int *color,
                              int *ctl, int order, const int *sceneStart, int *overflow, const int *rank) {
  if (ctl[1]) return;  // converged in an earlier launch
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && props[i].mass > 0) {
    // plain loads: a colour read before another thread's change in this
    // round is a past value, and that change makes the round count (only a
    // round with no change ends the relaxation), so no ordering is needed
    int *vc = color;
    const int base = sceneStart[props[i].scene];
    unsigned long long used[MAX_COLORS / 64] = {0, 0, 0, 0};
    // neighbours in chunks of 8: the chunk's indices are requested together,
    // then its colours, instead of one dependent pair of loads per neighbour
    constexpr int CH = 8;
    const int e1 = adjStart[i + 1];
    for (int e0 = adjStart[i]; e0 < e1; e0 += CH) {
      int j[CH], c[CH];
#pragma unroll
      for (int k = 0; k < CH; k++) j[k] = e0 + k < e1 ? adjOther[e0 + k] : -1;
#pragma unroll
      for (int k = 0; k < CH; k++) {
        if (j[k] >= 0 && !claimsBefore(j[k], i, order, base, rank)) j[k] = -1;
        c[k] = j[k] >= 0 ? vc[j[k]] : -1;
      }
#pragma unroll
      for (int k = 0; k < CH; k++)
        if (c[k] >= 0 && c[k] < MAX_COLORS) used[c[k] >> 6] |= 1ull << (c[k] & 63);
    }
    int c = MAX_COLORS;
    for (int w = 0; w < MAX_COLORS / 64; w++)
      if (~used[w]) { c = w * 64 + __builtin_ctzll(~used[w]); break; }
    if (c >= MAX_COLORS) { atomicOr(overflow, 4); c = MAX_COLORS - 1; }
    if (c != vc[i]) {
      vc[i] = c;
      atomicAdd(&ctl[0], 1);
    }
  }
  // the last block to finish closes the round: it counts it, marks the
  // colouring converged if nothing changed, and resets the counters
  // (ctl[2] counts the finished blocks)
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(&ctl[2], 1) == int(gridDim.x) - 1;
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    ctl[3] += 1;
    if (atomicAdd(&ctl[0], 0) == 0) ctl[1] = 1;
    ctl[0] = 0;
    ctl[2] = 0;
  }
}

// An uploaded colouring, checked each step: every dynamic body has a colour
// in range (else flag 1) and no force joins two bodies of one colour (else
// flag 2), which the colour-parallel sweep requires.
// ------------------------------------------------- colouring in one launch
// The same greedy colouring computed in one launch instead of relaxation
// rounds: a body takes the smallest colour no claims-before neighbour holds
// once every such neighbour holds its final colour. state[i] publishes body
// i's result, depth << 9 | colour (depth: 1 + the largest depth among its
// claims-before neighbours, the relaxation rounds it takes to settle), or
// DAG_OPEN until then; a static body is DAG_STATIC (it claims nothing).
constexpr int DAG_OPEN = -1, DAG_STATIC = -2;

__global__ void k_color_dag_init(int n, const BodyProp *props, int *state, int *ctl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) state[i] = props[i].mass > 0 ? DAG_OPEN : DAG_STATIC;
  if (i == 0) { ctl[0] = 0; ctl[2] = 0; ctl[3] = 0; ctl[6] = 0; }
}

// Waves take the dynamic bodies in claim order (corder), 64 at a time from
// one counter (ctl[2]). A body waits only for bodies before it in that order:
// those of earlier chunks, which waves already running hold, and those of its
// own chunk. No lane spins alone: each pass over the chunk colours every lane
// whose claims-before neighbours are all final, so a chunk finishes in at
// most as many passes as its longest chain, and the launch completes whatever
// the number of waves resident. A wave that has made no progress for spinCap
// passes stops and leaves its open bodies uncoloured (ctl[0] counts the
// coloured bodies; k_color_dag_close reports the colouring unfinished).
// state words are read and written with device-scope relaxed atomics: each
// carries its whole payload, so no other ordering is needed.
__global__ void __launch_bounds__(256) k_color_dag(int nDyn, const int *corder, const BodyProp *props,
                                                   const int *adjStart, const int *adjOther, int *color, int *state,
                                                   int order, const int *sceneStart, const int *rank, int *ctl,
                                                   int *overflow, int spinCap) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    int k0 = 0;
    if (lane == 0) k0 = atomicAdd(&ctl[2], 64);
    k0 = __shfl(k0, 0);
    if (k0 >= nDyn) break;
    const int k = k0 + lane;
    const int i = k < nDyn ? corder[k] : -1;
    bool open = i >= 0;
    const int base = open ? sceneStart[props[i].scene] : 0;
    const int e0 = open ? adjStart[i] : 0, e1 = open ? adjStart[i + 1] : 0;
    int idle = 0, coloured = 0, maxDepth = 0;
    // each lane's scan of its neighbours resumes at the first chunk not yet
    // all final (the colours and depths of a chunk's final neighbours are
    // merged again when the chunk is read again: the merges are idempotent)
    unsigned long long used[MAX_COLORS / 64] = {0, 0, 0, 0};
    int depth = 0, ePos = e0;
    while (__ballot(open) != 0ull) {
      bool progress = false;
      if (open) {
        bool ready = true;
        constexpr int CH = 8;
        for (; ePos < e1 && ready; ePos += CH) {
          int j[CH], st[CH];
#pragma unroll
          for (int q = 0; q < CH; q++) j[q] = ePos + q < e1 ? adjOther[ePos + q] : -1;
#pragma unroll
          for (int q = 0; q < CH; q++) {
            if (j[q] >= 0 && !claimsBefore(j[q], i, order, base, rank)) j[q] = -1;
            st[q] = j[q] >= 0 ? __hip_atomic_load(&state[j[q]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : DAG_STATIC;
          }
#pragma unroll
          for (int q = 0; q < CH; q++) {
            if (st[q] == DAG_OPEN) ready = false;
            if (st[q] < 0) continue;
            const int c = st[q] & 511;
            used[c >> 6] |= 1ull << (c & 63);
            depth = max(depth, st[q] >> 9);
          }
          if (!ready) break;  // this chunk again on the next pass
        }
        if (ready) {
          int c = MAX_COLORS;
          for (int w = 0; w < MAX_COLORS / 64; w++)
            if (~used[w]) { c = w * 64 + __builtin_ctzll(~used[w]); break; }
          if (c >= MAX_COLORS) { atomicOr(overflow, 4); c = MAX_COLORS - 1; }
          color[i] = c;
          const int d = min(depth + 1, (1 << 21) - 1);
          __hip_atomic_store(&state[i], (d << 9) | c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          maxDepth = max(maxDepth, d);
          coloured++;
          open = false;
          progress = true;
        }
      }
      idle = __ballot(progress) != 0ull ? 0 : idle + 1;
      if (idle > spinCap) {
        // no progress for spinCap passes: every wave stops (ctl[6]), and
        // k_color_dag_close reports the colouring unfinished
        if (lane == 0) __hip_atomic_store(&ctl[6], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      if (idle > 0 && __hip_atomic_load(&ctl[6], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    }
    if (coloured) {
      atomicAdd(&ctl[0], coloured);
      atomicMax(&ctl[3], maxDepth);
    }
    if (__hip_atomic_load(&ctl[6], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
  }
}

// ctl[1] = 1 when every dynamic body was coloured (else the relaxation rounds
// take over); ctl[3] holds the rounds the relaxation would take.
__global__ void k_color_dag_close(int nDyn, int *ctl) {
  ctl[1] = ctl[0] == nDyn && !ctl[6] ? 1 : 0;
  ctl[0] = 0;
  ctl[2] = 0;
  ctl[6] = 0;
}

__global__ void k_color_validate(int n, const BodyProp *props, const int *adjStart, const int *adjOther,
                                 const int *color, int *flag, int *overflow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(props[i].mass > 0)) return;
  const int c = color[i];
  int bad = (c < 0 || c >= MAX_COLORS) ? 1 : 0;
  for (int e = adjStart[i]; !bad && e < adjStart[i + 1]; e++) {
    const int j = adjOther[e];
    if (j >= 0 && props[j].mass > 0 && color[j] == c) bad = 2;
  }
  if (bad) {
    atomicOr(flag, bad);
    atomicOr(overflow, 16);  // the step falls back to the greedy colouring
  }
}

// An uploaded colouring in place of the relaxation: the colours of the
// dynamic bodies, -1 for the static ones; marks the colouring converged.
__global__ void k_color_install(int n, const BodyProp *props, const int *src, int *color, int *ctl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) ctl[1] = 1;
  if (i >= n) return;
  color[i] = props[i].mass > 0 ? src[i] : -1;
}


// ------------------------------------------------------ whole-step graphs
// A step replayed as one graph takes every count from device memory. pc:
// [PC_OLD] live manifolds the step starts with, [PC_NEW] new pairs taken,
// [PC_TOT] their sum, [PC_MISS] this step's misses (1: the relaxation had not
// reached its fixed point after the rounds launched, 2: more colours than
// launched, 4: a capacity was exceeded), [PC_SEEN] the colour count before
// the cap, [PC_SERIAL] steps written to the host ring, [PC_PAIRS] pairs found,
// [PC_WANT] manifolds asked for.
enum { PC_OLD, PC_NEW, PC_TOT, PC_MISS, PC_SEEN, PC_SERIAL, PC_PAIRS, PC_WANT, PC_N };

// After the broad phase: the narrow phase's counts, within the capacities
// (an excess is dropped and reported: bits 1 and 2 of the overflow word, as
// the reference's overflow counter reports them, layout.js:171-178).
__global__ void k_pipe_counts(int *scal, int *pc, int pairCap, int manCap) {
  const int nOld = scal[2], found = scal[0], fresh = scal[1];
  int nNew = fresh, miss = 0;
  if (found > pairCap) {
    nNew = min(nNew, pairCap);
    miss = 4;
    atomicOr(&scal[5], 1);
  }
  if (nOld + nNew > manCap) {
    nNew = manCap - nOld;
    miss = 4;
    atomicOr(&scal[5], 2);
  }
  pc[PC_OLD] = nOld;
  pc[PC_NEW] = nNew;
  pc[PC_TOT] = nOld + nNew;
  pc[PC_PAIRS] = found;
  pc[PC_WANT] = nOld + fresh;
  pc[PC_MISS] |= miss;
  scal[8] = 0;  // contacts, counted by the narrow phase
}

// After the relaxation and the cap: own* say whether the round count and the
// cap are the graph's own (then a miss asks the host for a larger graph) or
// the caller's limits (GpuBackend.colorRounds / encColors).
__global__ void k_pipe_color_check(const int *cmeta, int *pc, int *overflow, int cap, int ownRounds, int ownCap) {
  if (!cmeta[1]) {
    atomicOr(overflow, 32);
    if (ownRounds) pc[PC_MISS] |= 1;
  }
  if (ownCap && pc[PC_SEEN] > cap) pc[PC_MISS] |= 2;
}

// The step's counters into the host ring (pinned host memory, PIPE_STRIDE
// ints per step, slot = step serial mod slots): [0..13] counts, misses and
// the serial, [16..16+PIPE_COLORS] the first colour starts and
// [64..64+PIPE_COLORS) their heavy-body counts (the host sizes the next
// graphs' launches from them); this step's misses are reset.
constexpr int PIPE_STRIDE = 96, PIPE_COLORS = 32;
__global__ void k_pipe_stats(const int *scal, const int *cmeta, int *pc, int *ring, int slots) {
  const int serial = pc[PC_SERIAL];
  int *o = ring + (serial % slots) * PIPE_STRIDE;
  const int t = threadIdx.x;
  if (t <= PIPE_COLORS) o[16 + t] = cmeta[8 + t];
  if (t < PIPE_COLORS) o[64 + t] = cmeta[CMETA_HEAVY + t];
  if (t == 0) {
    o[0] = scal[0]; o[1] = scal[1]; o[2] = scal[2]; o[3] = scal[8]; o[4] = scal[5];
    o[5] = cmeta[4]; o[6] = cmeta[3]; o[7] = cmeta[1];
    o[8] = pc[PC_MISS]; o[9] = pc[PC_SEEN]; o[10] = pc[PC_PAIRS]; o[11] = pc[PC_WANT];
    o[12] = pc[PC_TOT]; o[13] = serial;
  }
  __syncthreads();
  if (t != 0) return;
  __threadfence_system();
  pc[PC_SERIAL] = serial + 1;
  pc[PC_MISS] = 0;
  pc[PC_SEEN] = 0;
}

// Flags in the y word of a contact work item (see k_items).
constexpr uint32_t ITEM_FIRST = 0x80000000u, ITEM_SPHERE = 0x40000000u, ITEM_BODY = 0x3fffffffu;

// Bucketing the dynamic bodies by colour without contended atomics: each
// 256-body block writes its colour histogram (hist[block][colour]), one block
// turns the histograms into per-block offsets and the colour starts, and each
// block then places its bodies (order within a colour and block is free: a
// colour's bodies share no force).
__global__ void k_color_hist(int n, const int *color, int *hist, int *numColors) {
  __shared__ int h[MAX_COLORS];
  __shared__ int top;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = threadIdx.x; k < MAX_COLORS; k += blockDim.x) h[k] = 0;
  if (threadIdx.x == 0) top = 0;
  __syncthreads();
  const int c = i < n ? color[i] : -1;
  if (c >= 0) {
    atomicAdd(&h[c], 1);
    atomicMax(&top, c + 1);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < MAX_COLORS; k += blockDim.x) hist[size_t(blockIdx.x) * MAX_COLORS + k] = h[k];
  if (threadIdx.x == 0 && top > 0) atomicMax(numColors, top);
}

// A cap on the colour count (GpuBackend.encColors, backend.js:1403-1430): a
// body whose colour is at or above the cap takes the last allowed colour,
// which may put two bodies of one force in one colour (reported in bit 32 of
// the overflow word).
// seen (or null) receives the colour count before the cap.
__global__ void k_color_cap(int n, int *color, int cap, int *overflow, int *seen) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = color[i];
  if (seen && c >= 0) atomicMax(seen, c + 1);
  if (c >= cap) {
    color[i] = cap - 1;
    atomicOr(overflow, 32);
  }
}

// One 256-thread block per colour: offs[block][colour], the colour's bodies in
// earlier blocks (each thread scans a contiguous run of blocks, then the runs
// are offset by a block-wide scan), and the colour's total.
__global__ void k_color_offsets(int nb, const int *hist, int *offs, int *ctotal) {
  __shared__ int sh[256];
  const int c = blockIdx.x, t = threadIdx.x;
  const int per = (nb + 255) / 256, b0 = t * per, b1 = min(nb, b0 + per);
  int run = 0;
  for (int b = b0; b < b1; b++) run += hist[size_t(b) * MAX_COLORS + c];
  sh[t] = run;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int u = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += u;
    __syncthreads();
  }
  int acc = sh[t] - run;
  for (int b = b0; b < b1; b++) {
    const int v = hist[size_t(b) * MAX_COLORS + c];
    offs[size_t(b) * MAX_COLORS + c] = acc;
    acc += v;
  }
  if (t == 255) ctotal[c] = sh[255];
}

// Colour starts: exclusive scan of the MAX_COLORS totals (one block).
// It also clears the heavy-body counts of each colour and sets the heavy
// threshold heavy[MAX_COLORS]: a body with more work items than
// max(hmin, twice the mean per dynamic body) is swept by a team twice as
// wide (k_primal); hmin = 0 turns the split off.
__global__ void k_color_starts(const int *ctotal, int *cstart, int *heavy, const int *itemTotal, int nDyn,
                               int hmin) {
  __shared__ int sh[MAX_COLORS];
  const int c = threadIdx.x;
  const int v = ctotal[c];
  heavy[c] = 0;
  if (c == 0) {
    const long mean2 = (2L * itemTotal[0] + nDyn - 1) / max(nDyn, 1);
    heavy[MAX_COLORS] = hmin > 0 ? int(max(long(hmin), mean2)) : 0;
  }
  sh[c] = v;
  __syncthreads();
  for (int off = 1; off < MAX_COLORS; off <<= 1) {
    const int u = c >= off ? sh[c - off] : 0;
    __syncthreads();
    sh[c] += u;
    __syncthreads();
  }
  cstart[c] = sh[c] - v;
  if (c == MAX_COLORS - 1) cstart[MAX_COLORS] = sh[c];
}

// Place the bodies: cbodies (the body list of each colour) and cwork, the
// same list with each body's work-item range and its number of joint items,
// which is all a primal team needs to find its work.
// Within a colour the bodies stay in index order (a body's place in its block
// is the number of lower-indexed bodies of its colour there), so the bodies of
// one scene of a batch form one run of each colour's list. A body above the
// heavy threshold is also listed in hwork, from its colour's start on, in no
// particular order (the bodies of one colour share no force, so the order in
// which they are swept does not change a bit). 256 threads.
__global__ void k_color_scatter(int n, const int *color, const int *offs, const int *cstart, int *cbodies,
                                const int *itemStart, const int *itemJoints, int4 *cwork, int *heavy,
                                int4 *hwork) {
  __shared__ int base[MAX_COLORS];
  __shared__ int col[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = threadIdx.x; k < MAX_COLORS; k += blockDim.x)
    base[k] = cstart[k] + offs[size_t(blockIdx.x) * MAX_COLORS + k];
  const int c = i < n ? color[i] : -1;
  col[threadIdx.x] = c;
  __syncthreads();
  if (c < 0) return;
  int rank = 0;
  for (int j = 0; j < int(threadIdx.x); j++) rank += col[j] == c;
  const int pos = base[c] + rank;
  const int4 w = make_int4(i, itemStart[i], itemStart[i + 1], itemJoints[i]);
  cbodies[pos] = i;
  cwork[pos] = w;
  const int H = heavy[MAX_COLORS];
  if (H > 0 && w.z - w.y > H) hwork[cstart[c] + atomicAdd(&heavy[c], 1)] = w;
}

// Mark each contact item whose body comes first in the colour order (the other
// body is static or of a later colour): that body's sweep applies the
// contact's dual update. A team of 8 lanes per body.
__global__ void k_item_flags(int n, const int *color, const int *itemStart, uint2 *items) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gt / 8, lane = gt % 8;
  if (i >= n) return;
  const int c = color[i];
  if (c < 0) return;
  for (int t = itemStart[i] + lane; t < itemStart[i + 1]; t += 8) {
    const uint2 w = items[t];
    if (w.x & 0x80000000u) continue;
    const int o = int(w.y & ITEM_BODY), co = color[o];
    // a capped colouring may leave two bodies of one force in one colour (the
    // conflicts the reference tolerates): the lower index is the first visitor
    items[t].y = (w.y & ~ITEM_FIRST) | ((co < 0 || co > c || (co == c && o > i)) ? ITEM_FIRST : 0u);
  }
}

// Per dynamic body: its contact/joint work items in the order the reference
// visits them (forces newest first, contacts in order; solver.js:443-448).
// Item: x = manifold slot (m << 4 | s << 1 | isA), or 0x80000000 | j << 1 |
// isA; y = the other body (-1 for a joint's world anchor), so that its delta
// record can be fetched together with the contact record. For a contact, y
// also carries two flags: a sphere is involved, and (set after colouring) this
// body is the contact's first visitor in a sweep, the one that applies the
// contact's dual update of the previous iteration.
template <int PASS>
__global__ void k_items(int n, int J, const BodyProp *props, const int *adjStart, const int *adjKey, const int *live,
                        const int4 *head, const JointDev *joints, int *itemCnt, const int *itemStart,
                        uint2 *items, int cap, int *overflow, int *itemJoints) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!(props[i].mass > 0)) {
    if (PASS == 0) itemCnt[i] = 0;
    if (PASS == 1) itemJoints[i] = 0;
    return;
  }
  int cnt = 0, nj = 0;
  const int base = PASS == 1 ? itemStart[i] : 0;
  if (PASS == 1 && itemStart[i + 1] > cap) {
    atomicOr(overflow, 8);
    itemJoints[i] = 0;
    return;
  }
  // forces newest first, in chunks of 8: the chunk's keys, then their
  // manifolds and heads, each in one round trip
  constexpr int CH = 8;
  const int e0 = adjStart[i];
  for (int eTop = adjStart[i + 1] - 1; eTop >= e0; eTop -= CH) {
    int key[CH], m[CH];
    int4 h[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) key[k] = eTop - k >= e0 ? adjKey[eTop - k] : -1;
#pragma unroll
    for (int k = 0; k < CH; k++) m[k] = key[k] >= J ? live[key[k] - J] : -1;
#pragma unroll
    for (int k = 0; k < CH; k++) h[k] = m[k] >= 0 ? head[m[k]] : make_int4(0, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < CH; k++) {
      if (key[k] < 0) break;
      if (key[k] < J) {
        if (PASS == 1) {
          const bool isA = joints[key[k]].a == i;
          items[base + cnt] = make_uint2(0x80000000u | (uint32_t(key[k]) << 1) | (isA ? 1u : 0u),
                                         uint32_t(isA ? joints[key[k]].b : joints[key[k]].a));
        }
        cnt++;
        nj++;
      } else {
        if (PASS == 1)
          for (int s = 0; s < h[k].z; s++)
            items[base + cnt + s] =
                make_uint2((uint32_t(m[k]) << 4) | (uint32_t(s) << 1) | (h[k].x == i ? 1u : 0u),
                           uint32_t(h[k].x == i ? h[k].y : h[k].x) | ((h[k].w & MF_SPHERE) ? ITEM_SPHERE : 0u));
        cnt += h[k].z;
      }
    }
  }
  if (PASS == 0) itemCnt[i] = cnt;
  if (PASS == 1) itemJoints[i] = nj;
}

// Algorithmic bytes of one primal visit of each dynamic body, summed per
// colour: own records, plus per manifold touched (head, basis, the other
// body's delta record) and per contact (item word, hot record).
// cbytes[c]: a sweep without the folded dual update; cbytes[MAX_COLORS + c]:
// what the folded update adds (k, lambda and the stick flag written by first
// visitors).
struct ByteModel { unsigned long long own, man, item, joint, dualw; };
__global__ void k_color_bytes(int n, int J, const BodyProp *props, const int *color, const int *adjStart,
                              const int *adjKey, const int *itemStart, const uint2 *items, ByteModel bm,
                              unsigned long long *cbytes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(props[i].mass > 0)) return;
  int nm = 0, nj = 0, nf = 0;
  for (int e = adjStart[i]; e < adjStart[i + 1]; e++) (adjKey[e] < J ? nj : nm)++;
  for (int t = itemStart[i]; t < itemStart[i + 1]; t++)
    if (!(items[t].x & 0x80000000u) && (items[t].y & ITEM_FIRST)) nf++;
  const int nc = itemStart[i + 1] - itemStart[i] - nj;
  const unsigned long long b = bm.own + bm.man * nm + bm.item * nc + bm.joint * nj;
  atomicAdd(&cbytes[color[i]], b);
  atomicAdd(&cbytes[MAX_COLORS + color[i]], bm.dualw * nf);
}

// ------------------------------------------------------------------ solve
// Body initialisation (solver.js:366-401), plus the per-body delta record
// dq = (p - p0, q (-) q0) that neighbours read.
template <class R>
__device__ __forceinline__ void bodyInitOne(int i, const BodyProp *props, double *X, double *X0, double *Y,
                                            const double *V, R *DQ, const StepParams &P) {
  const double dt = P.dt;
  const BodyProp &b = props[i];
  double *x = X + size_t(i) * XS, *x0 = X0 + size_t(i) * XS, *y = Y + size_t(i) * XS;
  const double *v = V + size_t(i) * VS;
  const double g[3] = {0, 0, P.gravity};
  double py[3] = {x[0], x[1], x[2]};
  for (int k = 0; k < 3; k++) py[k] += v[k] * dt;
  if (b.mass > 0)
    for (int k = 0; k < 3; k++) py[k] += g[k] * (dt * dt);
  double wdt[3] = {v[3] * dt, v[4] * dt, v[5] * dt}, qy[4];
  qint(qy, x + 3, wdt);
  const double accelZ = (v[2] - v[8]) / dt;
  const double accelExt = accelZ * sgn(P.gravity);
  double wgt = clampj(accelExt / xabs(P.gravity), 0.0, 1.0);
  if (!__builtin_isfinite(wgt)) wgt = 0.0;
  for (int k = 0; k < 7; k++) x0[k] = x[k];
  for (int k = 0; k < 3; k++) y[k] = py[k];
  for (int k = 0; k < 4; k++) y[3 + k] = qy[k];
  if (b.mass > 0) {
    for (int k = 0; k < 3; k++) x[k] += v[k] * dt;
    const double s = wgt * dt * dt;
    for (int k = 0; k < 3; k++) x[k] += g[k] * s;
    double q[4];
    qint(q, x + 3, wdt);
    for (int k = 0; k < 4; k++) x[3 + k] = q[k];
  }
  double dq[3];
  qsub(dq, x + 3, x0 + 3);
  R *d = DQ + size_t(i) * 8;
  d[0] = R(x[0] - x0[0]); d[1] = R(x[1] - x0[1]); d[2] = R(x[2] - x0[2]);
  d[3] = R(dq[0]); d[4] = R(dq[1]); d[5] = R(dq[2]);
  d[6] = R(0); d[7] = R(0);
}

template <class R>
__global__ void k_body_init(int n, const BodyProp *props, double *X, double *X0, double *Y, const double *V, R *DQ,
                            StepParams P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) bodyInitOne<R>(i, props, X, X0, Y, V, DQ, P);
}

struct SolveArgs {
  const BodyProp *props;
  double *X;
  const double *X0, *Y;
  const int *itemStart;
  const uint2 *items;
  const int4 *head;
  const JointDev *joints;
  int nBodies;
  int jointBodiesOnly;  // sweep only the bodies with joints (the exact team sweep leaves them)
};

template <class R>
__device__ __forceinline__ void loadRow(R *o, const R *p, int n) {
#pragma unroll
  for (int k = 0; k < n; k++) o[k] = p[k];
}

// Primal update of every body of one colour (solver.js:419-458): assemble
// H = M/dt^2 + sum J^T K J, f = M/dt^2 (x - y) + sum J^T lambda+, solve the
// 6x6 system in f64 and apply Eq. 21. TEAM lanes share a body: each stamps a
// strided share of its items and the team tree-reduces; TEAM = 1 with EXACT
// stamps in the reference's order onto the inertial terms. The colour's body
// range comes from device memory and teams stride over it, so any launch
// shape sweeps the whole colour (a graph captured at one step serves later
// steps whose colours have grown).
// Work-item words a lane fetches at once (one memory round trip for up to
// PRIMAL_BATCH of its items).
constexpr int PRIMAL_BATCH = 8;

// One colour of the primal sweep (solver.js:419-458): a team of TEAM lanes per
// body gathers its contacts and joints, reduces the 6x6 system and solves it.
// With dual set, the sweep also applies the previous iteration's dual update
// (manifold.js:447-491, evaluated with dualAlpha) to each contact whose first
// visitor this body is: both of its bodies are still where the previous
// iteration left them, so this is the update the separate dual pass would
// make, and the later visitor reads the updated k and lambda.
//
// In the exact mode (TEAM 1, f64) every term is added in the reference's
// order. Otherwise each lane sums its contacts in the record precision R (the
// contact terms are computed in R), the team reduces in f64, and lane 0 adds
// the inertial terms and the joints (cwork.w of them) in f64: joint rows can
// be stiff enough to swamp a light body's inertia in f32.
// One body's primal update (the body of the team loop below; TEAM lanes of
// one wave call it together for the same work entry w).
// GL: the GPU contact law (P.contactModel set), a template parameter so that
// the CPU law's sweep carries none of its code.
template <class R, bool EXACT, int TEAM, bool GL>
__device__ __forceinline__ void primalBody(const int4 w, const int lane, const SolveArgs &S, ManBuf<R> mb, R *DQ,
                                           const StepParams &P, double alpha, int dual, double dualAlpha) {
  using Acc = typename std::conditional<EXACT, double, R>::type;
  const double dt2 = P.dt * P.dt;
  const int i = w.x, b0 = w.y, b1 = w.z;
  const BodyProp &bp = S.props[i];
  double *x = S.X + size_t(i) * XS;
  const double *y = S.Y + size_t(i) * XS;
  const double *x0 = S.X0 + size_t(i) * XS;
  R dqi[8];
  vload<8>(dqi, DQ + size_t(i) * 8);
  // own pose and the step's start and inertial target, for lane 0's solve
  // (requested with the body's other records, so that they arrive during
  // the contact gathers; the exact mode, short of registers, asks after
  // them)
  double xs[7], ys[7], x0s[7];
  if constexpr (!EXACT)
    if (lane == 0)
      for (int k = 0; k < 7; k++) { xs[k] = x[k]; ys[k] = y[k]; x0s[k] = x0[k]; }
  // lane 0's inertial terms (f64, team modes): M/dt^2 (rotated inertia) and
  // M/dt^2 (x - y). Computed while the first contact records are in flight
  // (the wave would otherwise only wait for them), or after the gathers for
  // a body without contacts.
  double inMl = 0, inAng[6], inRl[3], inRa[3];
  bool inDone = false;
  auto inertia = [&]() {
    double MAng[9], Ma9[9], dq[3];
    worldMoment(MAng, xs + 3, bp.m0, bp.m1, bp.m2, !P.rotated || bp.shape == 1);
    const double idt2 = frcp(dt2);
    inMl = bp.mass * idt2;
    for (int k = 0; k < 9; k++) Ma9[k] = MAng[k] * idt2;
    qsub(dq, xs + 3, ys + 3);
    mulv(inRa, Ma9, dq);
    inAng[0] = Ma9[0]; inAng[1] = Ma9[3]; inAng[2] = Ma9[4];
    inAng[3] = Ma9[6]; inAng[4] = Ma9[7]; inAng[5] = Ma9[8];
    for (int k = 0; k < 3; k++) inRl[k] = inMl * (xs[k] - ys[k]);
    inDone = true;
  };
  Sys<Acc> sl;
  sl.zero();
  if constexpr (EXACT) {
    // inertial terms first, as the reference assembles them
    double MLin[9] = {bp.mass, 0, 0, 0, bp.mass, 0, 0, 0, bp.mass}, MAng[9], Ml[9], Ma[9];
    worldMoment(MAng, x + 3, bp.m0, bp.m1, bp.m2, !P.rotated || bp.shape == 1);
    for (int k = 0; k < 9; k++) { Ml[k] = MLin[k] / dt2; Ma[k] = MAng[k] / dt2; }
    double t[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]}, dq[3];
    mulv(sl.rl, Ml, t);
    qsub(dq, x + 3, y + 3);
    mulv(sl.ra, Ma, dq);
    sl.lin[0] = Ml[0]; sl.lin[1] = Ml[3]; sl.lin[2] = Ml[4]; sl.lin[3] = Ml[6]; sl.lin[4] = Ml[7]; sl.lin[5] = Ml[8];
    sl.ang[0] = Ma[0]; sl.ang[1] = Ma[3]; sl.ang[2] = Ma[4]; sl.ang[3] = Ma[6]; sl.ang[4] = Ma[7]; sl.ang[5] = Ma[8];
  }
  const R ralpha = R(alpha), rdalpha = R(dualAlpha);
  const R beta = R(P.betaLin), pmax = R(P.pmax), stickT = R(P.stick);
  const int lawP = GL ? contactLaw(1, P.postStab, alpha) : LAW_CPU;
  const int lawD = GL ? contactLaw(1, P.postStab, dualAlpha) : LAW_CPU;
  constexpr int IB = PRIMAL_BATCH;
  for (int t0 = b0 + lane; t0 < b1; t0 += TEAM * IB) {
    uint2 itw[IB];
#pragma unroll
    for (int k = 0; k < IB; k++) {
      const int tk = t0 + k * TEAM;
      itw[k] = tk < b1 ? S.items[tk] : make_uint2(0xffffffffu, 0u);
    }
#pragma unroll
    for (int k = 0; k < IB; k++) {
      const uint2 iw = itw[k];
      if (iw.x == 0xffffffffu) break;
      const bool isA = iw.x & 1u;
      if (iw.x & 0x80000000u) {
        if constexpr (EXACT) jointStamp(sl, S.joints[int((iw.x & 0x7fffffffu) >> 1)], isA, S.X, alpha);
        continue;
      }
      const size_t si = size_t(iw.x >> 4) * MAXC + ((iw.x >> 1) & 7u);
      R bas[BAS], hot[HOT], dqo[8];
      vload<BAS>(bas, mb.basis + size_t(iw.x >> 4) * BAS);
      vload<HOT>(hot, mb.hot + si * HOT);
      vload<8>(dqo, DQ + size_t(iw.y & ITEM_BODY) * 8);
      if constexpr (!EXACT && sizeof(R) == 4)  // f64 records leave no registers for it
        if (k == 0 && lane == 0 && !inDone) inertia();
      // the GPU contact law's cone depends on the stick flag
      int stk = GL ? mb.stick[si] : 0;
      if constexpr (!EXACT) {
        const R *dA = isA ? dqi : dqo, *dB = isA ? dqo : dqi;
        R rW[3];
#pragma unroll
        for (int q = 0; q < 3; q++) rW[q] = isA ? hot[HRA + q] : hot[HRB + q];
        FEval<R> e;
        if (dual && (iw.y & ITEM_FIRST)) {
          fevalC(bas, hot, dA, dB, rdalpha, e, lawD);
          fevalF(hot + HK, hot + HL, bas[BMU], e, lawD, bas[BMUD], stk);
          R kl[6];
          const int st = fdualUpdate(hot + HK, e, beta, pmax, stickT, iw.y & ITEM_SPHERE, kl, lawD);
          vstoreKL(mb.hot + si * HOT, kl);
          if (st >= 0) { mb.stick[si] = st; stk = st; }
          if (rdalpha != ralpha) fevalCAlpha(bas, hot, ralpha, e, lawP);
          fevalF(kl, kl + 3, bas[BMU], e, lawP, bas[BMUD], stk);
          fstamp<R, Acc>(sl, bas, isA ? R(1) : R(-1), rW, kl, e.F);
        } else {
          fevalC(bas, hot, dA, dB, ralpha, e, lawP);
          fevalF(hot + HK, hot + HL, bas[BMU], e, lawP, bas[BMUD], stk);
          fstamp<R, Acc>(sl, bas, isA ? R(1) : R(-1), rW, hot + HK, e.F);
        }
        continue;
      }
      R negb[9];
      for (int e = 0; e < 9; e++) negb[e] = -bas[e];
      const R *dA = isA ? dqi : dqo, *dB = isA ? dqo : dqi;
      Eval<R> e;
      if (dual && (iw.y & ITEM_FIRST)) {
        evalContact<R, EXACT>(bas, negb, bas[BMU], hot, dA, dB, rdalpha, e, lawD, stk);
        R kl[6];
        const int st = dualUpdate(hot + HK, e, beta, pmax, stickT, iw.y & ITEM_SPHERE, kl, lawD);
        vstoreKL(mb.hot + si * HOT, kl);
        if (st >= 0) { mb.stick[si] = st; stk = st; }
#pragma unroll
        for (int q = 0; q < 6; q++) hot[HK + q] = kl[q];  // k then lambda
        if (rdalpha != ralpha) evalC<R, EXACT>(bas, negb, hot, dA, dB, ralpha, e, lawP);
        evalF<R, EXACT>(hot + HK, hot + HL, bas[BMU], e, lawP, bas[BMUD], stk);
      } else {
        evalContact<R, EXACT>(bas, negb, bas[BMU], hot, dA, dB, ralpha, e, lawP, stk);
      }
      stampContact<R, Acc>(sl, isA ? bas : negb, isA ? e.jA : e.jB, hot + HK, e.F);
    }
  }
  if constexpr (EXACT)
    if (lane == 0)
      for (int k = 0; k < 7; k++) { xs[k] = x[k]; ys[k] = y[k]; x0s[k] = x0[k]; }
  Sys<double> s;
  if constexpr (TEAM > 1 && std::is_same<Acc, float>::value) {
    // f32 lane sums reduce in f32: the tree adds four roundings to the
    // lanes' own sums, at a third of the instructions of an f64 tree
#pragma unroll
    for (int k = 0; k < 6; k++) { sl.lin[k] = teamSumF<TEAM>(sl.lin[k]); sl.ang[k] = teamSumF<TEAM>(sl.ang[k]); }
#pragma unroll
    for (int k = 0; k < 9; k++) sl.cross[k] = teamSumF<TEAM>(sl.cross[k]);
#pragma unroll
    for (int k = 0; k < 3; k++) { sl.rl[k] = teamSumF<TEAM>(sl.rl[k]); sl.ra[k] = teamSumF<TEAM>(sl.ra[k]); }
  }
  for (int k = 0; k < 6; k++) { s.lin[k] = double(sl.lin[k]); s.ang[k] = double(sl.ang[k]); }
  for (int k = 0; k < 9; k++) s.cross[k] = double(sl.cross[k]);
  for (int k = 0; k < 3; k++) { s.rl[k] = double(sl.rl[k]); s.ra[k] = double(sl.ra[k]); }
  if constexpr (TEAM > 1 && !std::is_same<Acc, float>::value) {
#pragma unroll
    for (int k = 0; k < 6; k++) { s.lin[k] = teamSum<TEAM>(s.lin[k]); s.ang[k] = teamSum<TEAM>(s.ang[k]); }
#pragma unroll
    for (int k = 0; k < 9; k++) s.cross[k] = teamSum<TEAM>(s.cross[k]);
#pragma unroll
    for (int k = 0; k < 3; k++) { s.rl[k] = teamSum<TEAM>(s.rl[k]); s.ra[k] = teamSum<TEAM>(s.ra[k]); }
  }
  if constexpr (!EXACT) {
    // joint and spring rows in f64, spread over the team: each lane stamps
    // its share (lane 0 on top of the reduced contact terms, the others
    // from zero) and an f64 reduction brings them to lane 0 (w.w, the
    // body's joint-item count, is the same on every lane of the team)
    if (w.w > 0) {
      if (lane != 0) s.zero();
      for (int t = b0 + lane; t < b1; t += TEAM) {
        const uint2 itw = S.items[t];
        if (itw.x & 0x80000000u) jointStampF(s, S.joints[int((itw.x & 0x7fffffffu) >> 1)], itw.x & 1u, S.X, alpha);
      }
      if constexpr (TEAM > 1) {
#pragma unroll
        for (int k = 0; k < 6; k++) { s.lin[k] = teamSum<TEAM>(s.lin[k]); s.ang[k] = teamSum<TEAM>(s.ang[k]); }
#pragma unroll
        for (int k = 0; k < 9; k++) s.cross[k] = teamSum<TEAM>(s.cross[k]);
#pragma unroll
        for (int k = 0; k < 3; k++) { s.rl[k] = teamSum<TEAM>(s.rl[k]); s.ra[k] = teamSum<TEAM>(s.ra[k]); }
      }
    }
  }
  if (lane != 0) return;
  if constexpr (!EXACT) {
    if (!inDone) inertia();
    s.lin[0] += inMl; s.lin[2] += inMl; s.lin[5] += inMl;
    for (int k = 0; k < 6; k++) s.ang[k] += inAng[k];
    for (int k = 0; k < 3; k++) { s.rl[k] += inRl[k]; s.ra[k] += inRa[k]; }
  }
  double aLin[9] = {s.lin[0], 0, 0, s.lin[1], s.lin[2], 0, s.lin[3], s.lin[4], s.lin[5]};
  double aAng[9] = {s.ang[0], 0, 0, s.ang[1], s.ang[2], 0, s.ang[3], s.ang[4], s.ang[5]};
  double bl[3] = {-s.rl[0], -s.rl[1], -s.rl[2]}, ba[3] = {-s.ra[0], -s.ra[1], -s.ra[2]};
  double dxl[3], dxa[3];
  double xn[7];
  if constexpr (EXACT) {
    solve6(aLin, aAng, s.cross, bl, ba, dxl, dxa);
    qint(xn + 3, xs + 3, dxa);
  } else {
    solve6f(aLin, aAng, s.cross, bl, ba, dxl, dxa);
    qintf(xn + 3, xs + 3, dxa);
  }
  for (int k = 0; k < 3; k++) xn[k] = xs[k] + dxl[k];
  for (int k = 0; k < 7; k++) x[k] = xn[k];
  double dq[3];
  qsub(dq, xn + 3, x0s + 3);
  R *d = DQ + size_t(i) * 8;
  d[0] = R(xn[0] - x0s[0]); d[1] = R(xn[1] - x0s[1]); d[2] = R(xn[2] - x0s[2]);
  d[3] = R(dq[0]); d[4] = R(dq[1]); d[5] = R(dq[2]);
}

// One colour of the sweep, a team of TEAM lanes per body. In the team modes
// the colour's heavy bodies, those listed in hwork (their count in
// heavy[c], the threshold in heavy[MAX_COLORS]), are swept by the first
// wideBlocks blocks with teams of HEAVY_WIDEN x TEAM lanes, so that a body
// with several times the usual contacts does not hold up the launch (their
// blocks go first, as the longest work should start first); the other blocks
// sweep the other bodies. Without the split wideBlocks is 0 and the blocks
// sweep every body of the colour.
constexpr int HEAVY_WIDEN = 2;
template <class R, bool EXACT, int TEAM>
constexpr bool heavySplitTeam() { return !EXACT && TEAM >= 4 && TEAM * HEAVY_WIDEN <= 64; }

template <class R, bool EXACT, int TEAM, bool GL>
__global__ void __launch_bounds__(256, 2) k_primal(int c, const int *cstart, const int4 *cwork, const int *heavy,
                                                const int4 *hwork, int wideBlocks, SolveArgs S, ManBuf<R> mb,
                                                R *DQ, StepParams P, double alpha, int dual, double dualAlpha) {
  constexpr bool SPLIT = heavySplitTeam<R, EXACT, TEAM>();
  const int c0 = cstart[c];
  if constexpr (SPLIT) {
    if (int(blockIdx.x) < wideBlocks) {
      constexpr int TW = HEAVY_WIDEN * TEAM;
      const int gt = blockIdx.x * blockDim.x + threadIdx.x;
      const int lane = gt % TW, teams = wideBlocks * blockDim.x / TW;
      const int cnt = heavy[c];
      for (int team = gt / TW; team < cnt; team += teams) {
        const int4 w = hwork[c0 + team];
        if (S.jointBodiesOnly && w.w == 0) continue;
        primalBody<R, EXACT, TW, GL>(w, lane, S, mb, DQ, P, alpha, dual, dualAlpha);
      }
      return;
    }
  }
  const int H = SPLIT ? heavy[MAX_COLORS] : 0;
  const int gt = (int(blockIdx.x) - wideBlocks) * blockDim.x + threadIdx.x;
  const int lane = gt % TEAM, teams = (int(gridDim.x) - wideBlocks) * blockDim.x / TEAM;
  const int cnt = cstart[c + 1] - c0;
  for (int team = gt / TEAM; team < cnt; team += teams) {
    const int4 w = cwork[c0 + team];
    if (S.jointBodiesOnly && w.w == 0) continue;
    if (SPLIT && H > 0 && w.z - w.y > H) continue;  // swept by the wide teams
    primalBody<R, EXACT, TEAM, GL>(w, lane, S, mb, DQ, P, alpha, dual, dualAlpha);
  }
}

// ------------------------------------------------- exact mode, team-parallel
// A contact's 27 contributions to its body's system (manifold.js:411-445),
// each rounded as stampContact rounds it before adding: the Lin and Ang lower
// triangles, the full Cross block, then the linear and angular right-hand
// sides, in that order.
__device__ __forceinline__ void contactTerms(double *o, const double *jl, const double *ja, const double *k,
                                             const double *F) {
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const int r = LT_R[i], c = LT_C[i];
    o[i] = (jl[r] * k[0]) * jl[c] + (jl[3 + r] * k[1]) * jl[3 + c] + (jl[6 + r] * k[2]) * jl[6 + c];
    o[6 + i] = (ja[r] * k[0]) * ja[c] + (ja[3 + r] * k[1]) * ja[3 + c] + (ja[6 + r] * k[2]) * ja[6 + c];
  }
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      o[12 + r * 3 + c] = (ja[r] * k[0]) * jl[c] + (ja[3 + r] * k[1]) * jl[3 + c] + (ja[6 + r] * k[2]) * jl[6 + c];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    o[21 + r] = jl[r] * F[0] + jl[3 + r] * F[1] + jl[6 + r] * F[2];
    o[24 + r] = ja[r] * F[0] + ja[3 + r] * F[1] + ja[6 + r] * F[2];
  }
}

// Orders LDS accesses of the lanes of one wave (the team lives in one wave):
// the writes before it are seen by the reads after it.
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int XT_TEAM = 32;   // lanes per body
constexpr int XT_TERMS = 27;  // numbers of the 6x6 system
// doubles per item row in LDS (27 used): an odd count, so that the 32 lanes
// writing their rows (58-dword strides) hit 32 different bank pairs
constexpr int XT_ROW = 29;

// One colour of the exact sweep with a team of XT_TEAM lanes per body, giving
// the same bits as the one-lane exact sweep (k_primal<double, true, 1>): the
// reference adds each contact's contribution to the body's system in force
// order onto the inertial terms (solver.js:419-458), and floating-point
// addition is not associative, so only the additions are kept in that order.
//   1. Each lane evaluates one contact of a chunk of XT_TEAM (C, the clamped
//      force, the folded dual update of a first visitor: all per contact and
//      order-free) and writes its 27 rounded contributions to a row in LDS.
//   2. Lane e (< 27) adds entry e of the rows in item order onto its running
//      sum, which started from the inertial term e.
// Lane 0 then gathers the 27 sums, solves the 6x6 system in f64 with the
// reference's LDL^T and applies Eq. 21. A body with joints or springs
// (cwork.w > 0) is left to the one-lane sweep (skipped here): joint rows add
// into the system several times per force (jointStamp).
// One body's exact primal update by a team of XT_TEAM lanes of one wave,
// with L the team's LDS rows (XT_TEAM * XT_ROW doubles); w.w must be 0 (no
// joint items).
template <bool GL>
__device__ __forceinline__ void primalBodyXT(const int4 w, const int lane, double *L, const SolveArgs &S,
                                             ManBuf<double> mb, double *DQ, const StepParams &P, double alpha,
                                             int dual, double dualAlpha) {
  const double dt2 = P.dt * P.dt;
  const double beta = P.betaLin, pmax = P.pmax, stickT = P.stick;
  const int lawP = GL ? contactLaw(1, P.postStab, alpha) : LAW_CPU;
  const int lawD = GL ? contactLaw(1, P.postStab, dualAlpha) : LAW_CPU;
  const int i = w.x, b0 = w.y, b1 = w.z;
  const BodyProp &bp = S.props[i];
  double *x = S.X + size_t(i) * XS;
  const double *y = S.Y + size_t(i) * XS;
  const double *x0 = S.X0 + size_t(i) * XS;
  double dqi[8];
  vload<8>(dqi, DQ + size_t(i) * 8);
  // inertial terms, the reference's first contributions (lane 0 into row 0)
  if (lane == 0) {
    double MLin[9] = {bp.mass, 0, 0, 0, bp.mass, 0, 0, 0, bp.mass}, MAng[9], Ml[9], Ma[9];
    worldMoment(MAng, x + 3, bp.m0, bp.m1, bp.m2, !P.rotated || bp.shape == 1);
    for (int k = 0; k < 9; k++) { Ml[k] = MLin[k] / dt2; Ma[k] = MAng[k] / dt2; }
    double t[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]}, dq[3], rl[3], ra[3];
    mulv(rl, Ml, t);
    qsub(dq, x + 3, y + 3);
    mulv(ra, Ma, dq);
    const double lin[6] = {Ml[0], Ml[3], Ml[4], Ml[6], Ml[7], Ml[8]};
    const double ang[6] = {Ma[0], Ma[3], Ma[4], Ma[6], Ma[7], Ma[8]};
    for (int k = 0; k < 6; k++) { L[k] = lin[k]; L[6 + k] = ang[k]; }
    for (int k = 0; k < 9; k++) L[12 + k] = 0.0;
    for (int k = 0; k < 3; k++) { L[21 + k] = rl[k]; L[24 + k] = ra[k]; }
  }
  waveSync();
  double acc = lane < XT_TERMS ? L[lane] : 0.0;
  waveSync();
  for (int t0 = b0; t0 < b1; t0 += XT_TEAM) {
    const int t = t0 + lane;
    if (t < b1) {
      const uint2 iw = S.items[t];
      const bool isA = iw.x & 1u;
      const size_t si = size_t(iw.x >> 4) * MAXC + ((iw.x >> 1) & 7u);
      double bas[BAS], hot[HOT], dqo[8], negb[9];
      vload<BAS>(bas, mb.basis + size_t(iw.x >> 4) * BAS);
      vload<HOT>(hot, mb.hot + si * HOT);
      vload<8>(dqo, DQ + size_t(iw.y & ITEM_BODY) * 8);
      for (int e = 0; e < 9; e++) negb[e] = -bas[e];
      const double *dA = isA ? dqi : dqo, *dB = isA ? dqo : dqi;
      Eval<double> e;
      int stk = GL ? mb.stick[si] : 0;
      if (dual && (iw.y & ITEM_FIRST)) {
        evalContact<double, true>(bas, negb, bas[BMU], hot, dA, dB, dualAlpha, e, lawD, stk);
        double kl[6];
        const int st = dualUpdate(hot + HK, e, beta, pmax, stickT, iw.y & ITEM_SPHERE, kl, lawD);
        vstoreKL(mb.hot + si * HOT, kl);
        if (st >= 0) { mb.stick[si] = st; stk = st; }
#pragma unroll
        for (int q = 0; q < 6; q++) hot[HK + q] = kl[q];
        if (dualAlpha != alpha) evalC<double, true>(bas, negb, hot, dA, dB, alpha, e, lawP);
        evalF<double, true>(hot + HK, hot + HL, bas[BMU], e, lawP, bas[BMUD], stk);
      } else {
        evalContact<double, true>(bas, negb, bas[BMU], hot, dA, dB, alpha, e, lawP, stk);
      }
      double o[XT_TERMS];
      contactTerms(o, isA ? bas : negb, isA ? e.jA : e.jB, hot + HK, e.F);
      double *row = L + lane * XT_ROW;
#pragma unroll
      for (int k = 0; k < XT_TERMS; k++) row[k] = o[k];
    }
    waveSync();
    const int m = min(XT_TEAM, b1 - t0);
    if (lane < XT_TERMS) {
      // the rows in item order; eight reads issued at once, the additions
      // still one after the other
      const double *col = L + lane;
      int k = 0;
      for (; k + 8 <= m; k += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = col[(k + u) * XT_ROW];
#pragma unroll
        for (int u = 0; u < 8; u++) acc += v[u];
      }
      for (; k < m; k++) acc += col[k * XT_ROW];
    }
    waveSync();
  }
  // the 27 sums to lane 0
  if (lane < XT_TERMS) L[lane] = acc;
  waveSync();
  if (lane == 0) {
    Sys<double> s;
    for (int k = 0; k < 6; k++) { s.lin[k] = L[k]; s.ang[k] = L[6 + k]; }
    for (int k = 0; k < 9; k++) s.cross[k] = L[12 + k];
    for (int k = 0; k < 3; k++) { s.rl[k] = L[21 + k]; s.ra[k] = L[24 + k]; }
    double xs[7], x0s[7];
    for (int k = 0; k < 7; k++) { xs[k] = x[k]; x0s[k] = x0[k]; }
    double aLin[9] = {s.lin[0], 0, 0, s.lin[1], s.lin[2], 0, s.lin[3], s.lin[4], s.lin[5]};
    double aAng[9] = {s.ang[0], 0, 0, s.ang[1], s.ang[2], 0, s.ang[3], s.ang[4], s.ang[5]};
    double bl[3] = {-s.rl[0], -s.rl[1], -s.rl[2]}, ba[3] = {-s.ra[0], -s.ra[1], -s.ra[2]};
    double dxl[3], dxa[3], xn[7];
    solve6(aLin, aAng, s.cross, bl, ba, dxl, dxa);
    qint(xn + 3, xs + 3, dxa);
    for (int k = 0; k < 3; k++) xn[k] = xs[k] + dxl[k];
    for (int k = 0; k < 7; k++) x[k] = xn[k];
    double dq[3];
    qsub(dq, xn + 3, x0s + 3);
    double *d = DQ + size_t(i) * 8;
    d[0] = xn[0] - x0s[0]; d[1] = xn[1] - x0s[1]; d[2] = xn[2] - x0s[2];
    d[3] = dq[0]; d[4] = dq[1]; d[5] = dq[2];
  }
  waveSync();  // row 0 is the next body's
}

// BLOCK lanes per workgroup (BLOCK / XT_TEAM bodies at once).
template <bool GL, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_primal_xt(int c, const int *cstart, const int4 *cwork, SolveArgs S,
                                                   ManBuf<double> mb, double *DQ, StepParams P, double alpha,
                                                   int dual, double dualAlpha) {
  __shared__ double rows[BLOCK / XT_TEAM][XT_TEAM * XT_ROW];
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = gt % XT_TEAM, teams = gridDim.x * blockDim.x / XT_TEAM;
  double *L = rows[threadIdx.x / XT_TEAM];
  const int c0 = cstart[c], cnt = cstart[c + 1] - c0;
  for (int team = gt / XT_TEAM; team < cnt; team += teams) {
    const int4 w = cwork[c0 + team];
    if (w.w > 0) continue;  // joints: the one-lane sweep
    primalBodyXT<GL>(w, lane, L, S, mb, DQ, P, alpha, dual, dualAlpha);
  }
}

// BDF1 velocities (solver.js:476-488).
__device__ __forceinline__ void velocityOne(int i, const BodyProp *props, const double *X, const double *X0,
                                            double *V, double dt) {
  double *v = V + size_t(i) * VS;
  v[6] = v[0]; v[7] = v[1]; v[8] = v[2];
  if (props[i].mass > 0) {
    const double *x = X + size_t(i) * XS, *x0 = X0 + size_t(i) * XS;
    for (int k = 0; k < 3; k++) v[k] = (x[k] - x0[k]) / dt;
    double w[3];
    qsub(w, x + 3, x0 + 3);
    for (int k = 0; k < 3; k++) v[3 + k] = w[k] / dt;
  }
}

__global__ void k_velocity(int n, const BodyProp *props, const double *X, const double *X0, double *V, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) velocityOne(i, props, X, X0, V, dt);
}

// ------------------------------------------------ batched scenes, one launch
// The solve of a batch of independent scenes in one launch per step: one
// workgroup per scene runs its bodies' initialisation, every iteration's
// colours one after another with a workgroup barrier between them (bodies of
// one colour share no force, so a colour's bodies run in parallel, and the
// barrier orders the colours as the per-colour launches do), the joints' dual
// update after each iteration and the velocities (solver.js:366-489). A
// scene's bodies of colour c are one run of the colour's work list (k_color_
// scatter keeps index order), found by binary search. Each body's update is the
// per-colour sweep's own (primalBody, primalBodyXT), so the exact precision
// gives the same bits. The contact dual of the last iteration, which also
// records C, runs after this kernel as k_dual: nothing here reads it.
struct SceneSolve {
  const int *sceneStart;       // [S + 1] body ranges
  const int *sceneJointStart;  // [S + 1] joint ranges (joints grouped by scene)
  const int *cmeta;            // [4] colour count, [8 ...] colour starts
  const int4 *cwork;
  double *V;
  JointDev *joints;
  int iterations;
  double alpha;
};

// First index in [lo, hi) whose work entry's body is >= body.
__device__ __forceinline__ int firstBodyAtLeast(const int4 *cwork, int lo, int hi, int body) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cwork[mid].x < body) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// BLOCK: the workgroup (lanes per scene); BLOCK / TEAM bodies of a colour
// are updated at once (the colour tile).
template <class R, bool EXACT, int TEAM, bool GL, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_solve_scenes(SceneSolve Q, SolveArgs S, ManBuf<R> mb, R *DQ,
                                                     StepParams P) {
  constexpr int T = EXACT ? XT_TEAM : TEAM;
  __shared__ double rows[EXACT ? BLOCK / XT_TEAM : 1][EXACT ? XT_TEAM * XT_ROW : 1];
  __shared__ int rlo[MAX_COLORS], rhi[MAX_COLORS];
  const int sc = blockIdx.x, tid = threadIdx.x;
  const int n0 = Q.sceneStart[sc], n1 = Q.sceneStart[sc + 1];
  const int j0 = Q.sceneJointStart[sc], j1 = Q.sceneJointStart[sc + 1];
  const int lane = tid % T, team = tid / T;
  constexpr int nteams = BLOCK / T;
  double *X0 = const_cast<double *>(S.X0), *Y = const_cast<double *>(S.Y);
  for (int i = n0 + tid; i < n1; i += BLOCK) bodyInitOne<R>(i, S.props, S.X, X0, Y, Q.V, DQ, P);
  // the scene's run of each colour's work list, found once per step
  const int ncol = Q.cmeta[4];
  const int *cst = Q.cmeta + 8;
  for (int c = tid; c < ncol; c += BLOCK) {
    const int lo = firstBodyAtLeast(Q.cwork, cst[c], cst[c + 1], n0);
    rlo[c] = lo;
    rhi[c] = firstBodyAtLeast(Q.cwork, lo, cst[c + 1], n1);
  }
  __syncthreads();
  for (int it = 0; it < Q.iterations; it++) {
    // from the second iteration on, the first visitor of each contact folds
    // in the previous iteration's dual update
    const int fused = it >= 1 ? 1 : 0;
    for (int c = 0; c < ncol; c++) {
      const int lo = rlo[c], hi = rhi[c];
      if (lo == hi) continue;  // no body of this colour in the scene (uniform over the workgroup)
      for (int t = lo + team; t < hi; t += nteams) {
        const int4 w = Q.cwork[t];
        if constexpr (EXACT) {
          if (w.w > 0) {
            if (lane == 0) primalBody<double, true, 1, GL>(w, 0, S, mb, DQ, P, Q.alpha, fused, Q.alpha);
          } else {
            primalBodyXT<GL>(w, lane, rows[team], S, mb, DQ, P, Q.alpha, fused, Q.alpha);
          }
        } else {
          primalBody<R, false, TEAM, GL>(w, lane, S, mb, DQ, P, Q.alpha, fused, Q.alpha);
        }
      }
      __syncthreads();
    }
    for (int j = j0 + tid; j < j1; j += BLOCK) jointDualOne(Q.joints[j], S.X, P, Q.alpha);
    __syncthreads();
  }
  for (int i = n0 + tid; i < n1; i += BLOCK) velocityOne(i, S.props, S.X, S.X0, Q.V, P.dt);
}

// Dual update of every live manifold (manifold.js:447-491); the live count
// comes from device memory, threads stride over it.
template <class R, bool EXACT, bool GL>
__global__ void __launch_bounds__(128) k_dual(const int *nLivePtr, const int *live, ManBuf<R> mb, const R *DQ, StepParams P, double alpha,
                       int writeC) {
  const int nLive = *nLivePtr;
  const R ralpha = R(alpha), beta = R(P.betaLin), pmax = R(P.pmax), stickT = R(P.stick);
  const int law = GL ? contactLaw(1, P.postStab, alpha) : LAW_CPU;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nLive; k += gridDim.x * blockDim.x) {
    const int m = live[k];
    const int4 h = mb.head[m];
    R bas[BAS], negb[9], dqA[8], dqB[8];
    vload<BAS>(bas, mb.basis + size_t(m) * BAS);
    for (int e = 0; e < 9; e++) negb[e] = -bas[e];
    vload<8>(dqA, DQ + size_t(h.x) * 8);
    vload<8>(dqB, DQ + size_t(h.y) * 8);
    const bool hasSphere = h.w & MF_SPHERE;
    for (int sl = 0; sl < h.z; sl++) {
      const size_t si = size_t(m) * MAXC + sl;
      R *hp = mb.hot + si * HOT;
      R hot[HOT];
      vload<HOT>(hot, hp);
      Eval<R> e;
      evalContact<R, EXACT>(bas, negb, bas[BMU], hot, dqA, dqB, ralpha, e, law, GL ? mb.stick[si] : 0);
      if (writeC) {
        R *cd = mb.Cd + si * 4;
        cd[0] = e.C[0]; cd[1] = e.C[1]; cd[2] = e.C[2];
      }
      R kl[6];
      const int st = dualUpdate(hot + HK, e, beta, pmax, stickT, hasSphere, kl, law);
      if (st >= 0) mb.stick[si] = st;
      vstoreKL(hp, kl);
    }
  }
}


// ------------------------------------------------------------ repack
// A topology edit that keeps the survivors' state (the identity-preserving
// repack of backend.js:662-1301 over Solver.removeBodies, solver.js:261-300):
// body i of the new scene takes the device state of old body oldIndex[i], when
// it has one: pose, step-start pose, velocities and, for a body still
// dynamic, its last colour (the colouring's warm start).
__global__ void k_repack_bodies(int n, const int *oldIndex, const BodyProp *props, const double *oX,
                                const double *oX0, const double *oV, const int *oColor, double *X, double *X0,
                                double *V, int *color) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = oldIndex[i];
  if (o < 0) return;
  for (int k = 0; k < XS; k++) {
    X[size_t(i) * XS + k] = oX[size_t(o) * XS + k];
    X0[size_t(i) * XS + k] = oX0[size_t(o) * XS + k];
  }
  for (int k = 0; k < VS; k++) V[size_t(i) * VS + k] = oV[size_t(o) * VS + k];
  if (props[i].mass > 0 && oColor[o] >= 0) color[i] = oColor[o];
}

// Whether each live manifold survives: both its bodies do.
__global__ void k_repack_keep(int nLive, const int *live, const int4 *head, const int *remap, int *keep) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nLive) return;
  const int4 h = head[live[t]];
  keep[t] = remap[h.x] >= 0 && remap[h.y] >= 0 ? 1 : 0;
}

// One wave per live manifold: a survivor's whole record (head with its bodies
// renumbered, basis, contact slots with their warm-start state) goes to slot
// keepPos[t] of the fresh store, in force-list order.
template <class R>
__global__ void __launch_bounds__(64) k_repack_manifolds(const int *live, const int *keep, const int *keepPos,
                                                         const int *remap, ManBuf<R> src, ManBuf<R> dst,
                                                         int *dstLive) {
  const int t = blockIdx.x, lane = threadIdx.x;
  if (!keep[t]) return;
  const size_t m = size_t(live[t]), p = size_t(keepPos[t]);
  if (lane == 0) {
    const int4 h = src.head[m];
    dst.head[p] = make_int4(remap[h.x], remap[h.y], h.z, h.w);
    dstLive[p] = int(p);
  }
  for (int k = lane; k < BAS; k += 64) dst.basis[p * BAS + k] = src.basis[m * BAS + k];
  for (int k = lane; k < MAXC * HOT; k += 64) dst.hot[p * MAXC * HOT + k] = src.hot[m * MAXC * HOT + k];
  for (int k = lane; k < MAXC * 4; k += 64) dst.Cd[p * MAXC * 4 + k] = src.Cd[m * MAXC * 4 + k];
  for (int k = lane; k < MAXC * 6; k += 64) dst.anchor[p * MAXC * 6 + k] = src.anchor[m * MAXC * 6 + k];
  for (int k = lane; k < MAXC; k += 64) {
    dst.feat[p * MAXC + k] = src.feat[m * MAXC + k];
    dst.stick[p * MAXC + k] = src.stick[m * MAXC + k];
  }
}

// Poses for drawing (renderer.js:1327-1383): each body at a fraction a of the
// last step, blended from the pose x_t at the step's start: the position
// lerped, the orientation nlerped along the shorter arc and renormalised, in
// f64, then rounded to f32. With lerp off (a >= 1) the current pose is copied.
// Out: per body two float4, {p, 0} and {q}. One thread per body; the 64 B pose
// records are read as four 16 B loads each, the output as two 16 B stores.
__global__ void k_export_poses(int n, double a, int lerp, const double2 *__restrict__ X,
                               const double2 *__restrict__ X0, float4 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int R2 = XS / 2;  // double2 per record
  const double2 p01 = X[size_t(i) * R2], p2q0 = X[size_t(i) * R2 + 1], q12 = X[size_t(i) * R2 + 2],
                q3 = X[size_t(i) * R2 + 3];
  double p[3] = {p01.x, p01.y, p2q0.x}, q[4] = {p2q0.y, q12.x, q12.y, q3.x};
  if (lerp) {
    const double2 o01 = X0[size_t(i) * R2], o2q0 = X0[size_t(i) * R2 + 1], oq12 = X0[size_t(i) * R2 + 2],
                  oq3 = X0[size_t(i) * R2 + 3];
    const double p0[3] = {o01.x, o01.y, o2q0.x}, q0[4] = {o2q0.y, oq12.x, oq12.y, oq3.x};
    for (int k = 0; k < 3; k++) p[k] = p0[k] + (p[k] - p0[k]) * a;
    const double d = q0[0] * q[0] + q0[1] * q[1] + q0[2] * q[2] + q0[3] * q[3];
    const double sg = d < 0 ? -1.0 : 1.0;
    for (int k = 0; k < 4; k++) q[k] = q0[k] * sg + (q[k] - q0[k] * sg) * a;
    const double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const double inv = 1.0 / (len != 0.0 ? len : 1.0);
    for (int k = 0; k < 4; k++) q[k] *= inv;
  }
  out[2 * size_t(i)] = make_float4(float(p[0]), float(p[1]), float(p[2]), 0.0f);
  out[2 * size_t(i) + 1] = make_float4(float(q[0]), float(q[1]), float(q[2]), float(q[3]));
}

}  // namespace avbd
