// avbd_api.hip — context, step orchestration and the C ABI of include/avbd.h.
//
// One context owns one HIP stream and every device buffer. avbd_step runs the
// kernels of avbd_kernels.hpp in the order of the reference's Solver.step
// (solver.js:307-493), with the primal sweep scheduled by the greedy colouring
// of the force graph (coloring.js:43-118) taken after force initialisation.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/avbd.h"
#include "avbd_adjoint.hpp"
#include "avbd_kernels.hpp"

using namespace avbd;

// Staging for avbd_export_poses_static: the poses of up to
// AVBD_STATIC_POSE_BODIES bodies (two float4 each), in device memory that
// every GPU the library runs on holds from module load until the process
// ends, so that the entry point allocates nothing. One copy per device, shared
// by every call and context on it; it keeps the last call's contents. The
// name is only a stand-in in host code: kernels and copies use the device
// address looked up per device (staticPoseScratch), never the name.
static __device__ __align__(256) float4 scratch[2 * AVBD_STATIC_POSE_BODIES];

namespace {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ApiError : std::runtime_error {
  int code;
  ApiError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                                                  \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess)                                                                            \
      throw HipError(std::string(#x) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" +        \
                     std::to_string(__LINE__) + ")");                                                \
  } while (0)

inline int blocks(long n, int b = 256) { return int((n + b - 1) / b); }

template <class T> struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  void reserve(size_t m) {
    if (m <= n && p) return;
    if (p) HIPCHECK(hipFree(p));
    p = nullptr;
    n = std::max<size_t>(m, 1);
    HIPCHECK(hipMalloc(&p, n * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // Device copy of the first count elements of src (grows as needed).
  void copyFrom(const T *src, size_t count, hipStream_t st) {
    reserve(count);
    if (count) HIPCHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyDeviceToDevice, st));
  }
};

struct ManStore {
  DBuf<int4> head;
  DBuf<char> basis, hot, Cd;
  DBuf<double> anchor;
  DBuf<uint32_t> feat;
  DBuf<int32_t> stick;
  void reserve(size_t m, size_t rsize) {
    head.reserve(m);
    basis.reserve(m * BAS * rsize);
    hot.reserve(m * MAXC * HOT * rsize);
    Cd.reserve(m * MAXC * 4 * rsize);
    anchor.reserve(m * MAXC * 6);
    feat.reserve(m * MAXC);
    stick.reserve(m * MAXC);
  }
  template <class R> ManBuf<R> view() {
    return ManBuf<R>{head.p, (R *)basis.p, (R *)hot.p, anchor.p, feat.p, stick.p, (R *)Cd.p};
  }
  void release() {
    head.release(); basis.release(); hot.release(); Cd.release(); anchor.release(); feat.release(); stick.release();
  }
};

struct Phase {
  hipEvent_t ev[AVBD_NUM_PHASES];
};

// Per device: the address of `scratch` on that device, looked up once with
// the device current (each GPU holds its own copy at its own address).
std::mutex scratchMutex;
std::map<int, float4 *> scratchAddr;

float4 *staticPoseScratch(int device) {
  std::lock_guard<std::mutex> lock(scratchMutex);
  auto it = scratchAddr.find(device);
  if (it != scratchAddr.end()) return it->second;
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  (void)hipSetDevice(prev);
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return nullptr;
  }
  scratchAddr[device] = static_cast<float4 *>(ptr);
  return static_cast<float4 *>(ptr);
}

}  // namespace

struct avbd_ctx {
  std::string err;
  int device = 0, prec = AVBD_PREC_MIXED, team = 16;
  // exact precision: a team of lanes per body with the additions in the
  // reference's order (k_primal_xt), or one lane per body (config team 1)
  bool exactTeam = true;
  // workgroup of the exact team sweep (avbd_set_launch_shape): 64 lanes, two
  // bodies, by default (the 51k pyramid: 8.99 ms per step against 9.38 at 128
  // and 9.85 at 256)
  int exactBlock = 64;
  int simds = 1024;  // SIMDs of the device (4 per CU), read at creation
  int manCapCfg = 0, pairCapCfg = 0;
  int needPairs = 0, needMan = 0;  // sizes an overflow asked for, for the next pack
  hipStream_t stream = nullptr;
  avbd_params prm{};
  bool packed = false;

  // sizes
  int N = 0, J = 0, S = 1, nLarge = 0;
  int manCap = 0, pairCap = 0;
  uint32_t tableMask = 0;
  double invh = 1.0;
  int64_t steps = 0;

  // bodies
  DBuf<BodyProp> props;
  DBuf<double> X, X0, Y, V;
  DBuf<char> DQ;
  DBuf<float4> poseStage;  // avbd_export_poses to host memory
  float4 *poseScratch = nullptr;  // this device's `scratch` (avbd_export_poses_static)
  std::vector<double> anchorStage;  // avbd_set_joint_anchor's host copies
  std::vector<int> hjointKind;      // AVBD_JOINT_* of each packed joint
  std::vector<BodyProp> hprops;
  // the GPU contact model's material: per body as given (static, dynamic
  // friction, restitution, contact and rest offset, NaN = the solver's), and
  // resolved on the device with the contact offsets the pair test reads
  std::vector<double> hmatRaw;
  DBuf<BodyMat> mat;
  DBuf<double> skinOff;
  double rmaxSmall = 0;  // largest radius of a body on the broad-phase grid

  // broad phase
  DBuf<int4> cellOf;
  DBuf<int> bucketCnt, bucketStart, bucketCursor, entries, bpScratch;
  DBuf<int4> entryCell;
  DBuf<int> sceneLargeStart, sceneLarges, sceneStart, largeList;
  // batches: the solve as one launch with a workgroup per scene
  DBuf<int> sceneJointStart;
  int sceneSolveCfg = 0;     // avbd_config.scene_solve
  bool sceneSolveOK = false; // the packed batch qualifies (2+ scenes, small, joints grouped)
  int sceneTeam = 0;         // lanes per body in it (0: from team)
  int sceneBlock = 0;        // lanes per scene (0: 256)
  DBuf<int> countRev, startRev, countNewRev, startNewRev;
  DBuf<int2> pairs, newPairs;
  int nPairs = 0, nNew = 0;

  // manifolds (double-buffered)
  ManStore store[2];
  DBuf<int> live[2];
  int cur = 0, nLive = 0;
  DBuf<int> keep, keepPos;

  // joints
  DBuf<JointDev> joints;

  // topology
  DBuf<int> deg, adjStart, adjCursor, adjKey, adjOther;
  DBuf<int> adjKeyFill, adjOtherFill, adjOwner;  // lists as filled, before the key order
  DBuf<int> color[2];
  int colorCur = 0;
  DBuf<int> cstart, cbodies, cmeta;
  DBuf<int4> cwork, hwork;  // each colour's bodies; its heavy bodies (k_color_scatter)
  std::vector<int> hcstart, hheavy, hmeta = std::vector<int>(CMETA_N);
  // the heavy-body split of the team sweeps (k_primal): on unless a caller
  // measures without it
  bool heavySplit = true;
  int colorBatch = 16;  // relaxation rounds launched between convergence checks
  // limits a host may set (GpuBackend.colorRounds / encColors): relaxation
  // rounds per step and colours per step; 0 = none (the exact greedy fixed point)
  int roundCap = 0, colorCap = 0;
  // The colour array is the greedy fixed point of the force graph as it stood
  // at the last colouring, in claim order colorFor (-1: no valid colouring).
  int colorFor = -1;
  // verification hooks: a caller's claim order (rank per body) and a caller's
  // colouring for the following steps
  bool useRank = false, useUpload = false;
  DBuf<int> claimRank, uploadColor;
  // the colouring in one launch (k_color_dag): the dynamic bodies in claim
  // order (built on the host for the order it was built for), each body's
  // published result, and the host's copies it is built from
  DBuf<int> corder, dagState;
  int corderFor = -1, nDynOrder = 0;
  std::vector<int> hsceneStart, hrank;
  bool dagColoring = true;  // off after a step reported it unfinished
  int nDynamic = 0;

  // solve-loop graphs, keyed by colour count and manifold buffer parity
  // (the captured graph is kept with its executable until both are dropped)
  std::map<long long, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
  bool graphCapturing = false;
  int graphsCfg = 0;
  // Whole steps as replayed graphs (avbd_config.graphs = AVBD_GRAPHS_STEP):
  // no host read inside a step. The counts the host keeps (live manifolds,
  // colours, overflow) reach it through a ring of pinned host slots that each
  // step's last kernel fills; a slot is read when it comes round again (at
  // most PIPE_SLOTS steps in flight) or when a call needs the state.
  static constexpr int PIPE_SLOTS = 4;
  DBuf<int> pc;                  // the step's device counts (PC_*)
  int *pipeRing = nullptr;       // pinned host ring, PIPE_SLOTS x PIPE_STRIDE ints
  int *pipeRingDev = nullptr;    // its device address
  hipEvent_t pipeEv[PIPE_SLOTS] = {};
  int64_t pipeSerial = 0, pipeTaken = 0;  // steps enqueued, ring slots read
  bool pipeReady = false;        // a step with host reads has run since the pack (shapes known)
  bool pipeUnsettled = false;    // the host's counts lag the device's
  bool pipeCapacity = false;     // a graph step exceeded a capacity: the next step call fails
  int pipeColors = 0, pipeRounds = 0;  // colour launches and relaxation rounds per graph (0: derived)
  // the counts a step graph's launch shapes are sized for (bodies per colour,
  // manifolds); a step that outgrows them has the graphs captured again
  std::vector<long> pipeEst, pipeEstH;  // colour and heavy-body counts the step graphs are shaped for
  long pipeTotEst = 0;
  bool pipeCapturing = false;
  // autotuner
  bool tuneRequest = false;
  float tuneMs[6] = {0};
  // every trial of the last autotune: [team 1, 4, 8, 16, 32, 64][workgroup
  // 64, 128, 256, 512] in ms (0: not tried)
  float tuneGrid[6][4] = {};
  int sweepBlock = 256;  // workgroup of the per-colour sweep (tuned with the team)
  DBuf<char> Xbk, DQbk;
  int numColors = 0, colorRounds = 0;
  DBuf<int> itemCnt, itemStart, itemJoints, chist, coffs, ctotal;
  DBuf<uint2> items;

  // ------------------------------------------------------------- adjoint
  // One recorded step (what the reverse kernels read) and a checkpoint of the
  // state one step hands to the next.
  struct TapeStep {
    int nTot = 0, nLive = 0, numColors = 0;
    DBuf<int4> head, cwork;
    DBuf<double> basis, anchor, hot, X, X0, Y, V;
    DBuf<int> src, live, color, cstart;
    DBuf<uint2> items;
    void release() {
      head.release(); cwork.release(); basis.release(); anchor.release(); hot.release(); X.release();
      X0.release(); Y.release(); V.release(); src.release(); live.release(); color.release(); cstart.release();
      items.release();
    }
  };
  struct Checkpoint {
    DBuf<double> X, V, anchor;
    DBuf<int4> head;
    DBuf<char> basis, hot, Cd;
    DBuf<uint32_t> feat;
    DBuf<int32_t> stick;
    DBuf<int> live, adjStart, adjKey, adjOther, color;
    int cur = 0, nLive = 0, lastNTot = 0;
    int64_t steps = 0;
    void release() {
      X.release(); V.release(); anchor.release(); head.release(); basis.release(); hot.release(); Cd.release();
      feat.release(); stick.release(); live.release(); adjStart.release(); adjKey.release(); adjOther.release();
      color.release();
    }
  };
  std::vector<TapeStep> tape;
  // The backward with segment re-runs reverses one segment on rstream while
  // the segment before it is re-run on stream, into the other half of a
  // double-buffered segment tape (default priority, like stream).
  hipStream_t rstream = nullptr;
  hipEvent_t evRerun = nullptr, evRev[2] = {nullptr, nullptr};
  TapeStep *rec = nullptr;  // the step being recorded, if any
  DBuf<int> srcSlot;        // per new contact: matched old slot | kept-anchor bit, or -1
  int lastNTot = 0;
  std::vector<Checkpoint> ckpts;
  Checkpoint endState;
  int adjSteps = 0, adjEvery = 0, adjSmallMax = 64;
  // Body state was overwritten since the last pack (avbd_write_state): the
  // velocity and the previous velocity may then differ, and the adjoint's
  // initial-velocity gradient (their sum) would be wrong.
  bool stateWritten = false;
  // A call that changes what a recorded forward stepped through (parameters,
  // state, schedule, anchors) or takes further steps drops the recording:
  // the backward then reports AVBD_E_STATE instead of reversing a trajectory
  // the context no longer holds.
  void dropAdjoint() { adjSteps = 0; }
  long adjTapeBudget = 4096L << 20;  // bytes the whole forward tape may take
  bool tapeResident = false;         // the last adjoint forward kept every step's tape
  DBuf<double> aXb, aX0b, aYb, aVb, aVbS, aLK[2], aC0b, aRWb, aMUb, aMu;

  // small device scalars: 0 pairs, 1 new, 2 live, 3 changed, 4 numColors,
  // 5 overflow, 6 items, 7 adj total
  DBuf<int> scal;
  int *hscal = nullptr;  // pinned mirror

  // scan temporaries per level
  std::vector<DBuf<int>> scanTmp = std::vector<DBuf<int>>(16);

  // profiling
  bool profiling = false;
  std::vector<hipEvent_t> evs;
  double profMs[AVBD_NUM_PHASES] = {0};
  int profSteps = 0;
  // per-kernel event pairs awaiting harvest, and the totals
  struct Pending { int kid; hipEvent_t a, b; double bytes; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> evPool;
  double kMs[AVBD_NUM_KERNELS] = {0}, kBytes[AVBD_NUM_KERNELS] = {0};
  int64_t kLaunches[AVBD_NUM_KERNELS] = {0};
  DBuf<unsigned long long> cbytes;
  std::vector<unsigned long long> hcbytes;
  int contactsTotal = 0;

  hipEvent_t takeEvent() {
    if (evPool.empty()) {
      hipEvent_t e;
      HIPCHECK(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = evPool.back();
    evPool.pop_back();
    return e;
  }
  // Launch a kernel on the stream. When profiling, the launch carries a pair of
  // events that the dispatch itself stamps at its begin and end
  // (hipExtLaunchKernel), so the time is the kernel's own duration, the same
  // interval a kernel trace reports, without event packets in between.
  template <class... KArgs, class... Args>
  void timedLaunch(int kid, double bytes, void (*kernel)(KArgs...), dim3 grid, dim3 block, Args... args) {
    if (!profiling) {
      hipLaunchKernelGGL(kernel, grid, block, 0, stream, args...);
      return;
    }
    Pending p{kid, takeEvent(), takeEvent(), bytes};
    hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, p.a, p.b, 0, args...);
    HIPCHECK(hipGetLastError());
    pending.push_back(p);
  }
  void harvest() {
    for (auto &p : pending) {
      HIPCHECK(hipEventSynchronize(p.b));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, p.a, p.b));
      kMs[p.kid] += ms;
      kBytes[p.kid] += p.bytes;
      kLaunches[p.kid]++;
      evPool.push_back(p.a);
      evPool.push_back(p.b);
    }
    pending.clear();
  }
  // Algorithmic bytes per unit of work for the record layout in use.
  ByteModel primalModel() const {
    const unsigned long long rs = rsize();
    ByteModel m;
    m.own = 7 * 8 * 3 + 40 + 16 + 6 * rs + 7 * 8 + 6 * rs;  // X, Y, X0, props, work entry, delta; X and delta written
    m.man = 10 * rs + 6 * rs;                               // basis + mu, other body's delta
    m.item = 8 + 15 * rs;                                   // item, hot contact record
    m.joint = sizeof(JointDev) + 2 * 7 * 8 + 4;    // joint record, both bodies' poses
    m.dualw = 6 * rs + 4;                           // k, lambda and stick written by a first visitor
    return m;
  }
  double dualBytes() const {
    const double rs = double(rsize());
    return double(nLive) * (4 + 16 + 10 * rs + 12 * rs) + double(contactsTotal) * (15 * rs + 6 * rs + 4);
  }

  ~avbd_ctx() {
    releaseAll();
    for (auto &t : tape) t.release();
    tape.clear();
  }

  void releaseAll() {
    dropGraphs();
    Xbk.release(); DQbk.release();
    props.release(); X.release(); X0.release(); Y.release(); V.release(); DQ.release(); poseStage.release();
    mat.release(); skinOff.release();
    cbytes.release();
    for (auto &p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    pending.clear();
    for (auto &e : evPool) (void)hipEventDestroy(e);
    evPool.clear();
    cellOf.release(); bucketCnt.release(); bucketStart.release(); bucketCursor.release(); entries.release();
    bpScratch.release(); entryCell.release();
    sceneLargeStart.release(); sceneLarges.release(); sceneStart.release(); largeList.release();
    sceneJointStart.release();
    countRev.release(); startRev.release(); countNewRev.release(); startNewRev.release(); cmeta.release(); pairs.release(); newPairs.release();
    store[0].release(); store[1].release(); live[0].release(); live[1].release(); keep.release(); keepPos.release();
    joints.release(); deg.release(); adjStart.release(); adjCursor.release(); adjKey.release(); adjOther.release();
    adjKeyFill.release(); adjOtherFill.release(); adjOwner.release();
    color[0].release(); color[1].release(); cstart.release(); cbodies.release(); cwork.release(); hwork.release();
    itemCnt.release(); itemStart.release(); items.release(); scal.release();
    itemJoints.release(); chist.release(); coffs.release(); ctotal.release();
    claimRank.release(); uploadColor.release();
    corder.release(); dagState.release(); corderFor = -1; nDynOrder = 0; dagColoring = true;
    // the tape's buffers are kept for the next adjoint forward (an
    // optimisation loop packs and records the same scene again); only the
    // record of which steps they hold is dropped
    tapeResident = false;
    adjSteps = 0;
    for (auto &c : ckpts) c.release();
    ckpts.clear();
    endState.release();
    srcSlot.release();
    aXb.release(); aX0b.release(); aYb.release(); aVb.release(); aVbS.release(); aLK[0].release(); aLK[1].release();
    aC0b.release(); aRWb.release(); aMUb.release(); aMu.release();
    for (auto &b : scanTmp) b.release();
    for (auto &e : evs) (void)hipEventDestroy(e);
    evs.clear();
    if (hscal) (void)hipHostFree(hscal);
    hscal = nullptr;
    if (pipeRing) (void)hipHostFree(pipeRing);
    pipeRing = pipeRingDev = nullptr;
    for (auto &e : pipeEv) {
      if (e) (void)hipEventDestroy(e);
      e = nullptr;
    }
    pipeUnsettled = false;
    pc.release();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    if (rstream) (void)hipStreamDestroy(rstream);
    rstream = nullptr;
    for (hipEvent_t *e : {&evRerun, &evRev[0], &evRev[1]}) {
      if (*e) (void)hipEventDestroy(*e);
      *e = nullptr;
    }
  }

  size_t rsize() const { return prec == AVBD_PREC_MIXED ? sizeof(float) : sizeof(double); }
  // the GPU contact law: its kernels are separate instantiations
  bool gl() const { return prm.contact_model != AVBD_CONTACTS_CPU; }

  StepParams sp() const {
    StepParams P;
    P.dt = prm.dt; P.gravity = prm.gravity; P.betaLin = prm.beta_lin; P.betaAng = prm.beta_ang;
    P.alpha = prm.alpha; P.gamma = prm.gamma; P.pmin = prm.penalty_min; P.pmax = prm.penalty_max;
    P.margin = prm.collision_margin; P.stick = prm.stick_threshold; P.iterations = prm.iterations;
    P.postStab = prm.post_stabilize; P.rotated = prm.rotated_inertia;
    P.contactModel = prm.contact_model;
    P.restThreshold = prm.restitution_threshold;
    P.persistence = prm.contact_persistence;
    return P;
  }

  // ---------------------------------------------------------------- scan
  // Exclusive scan of n ints; *total (device) receives the sum.
  // Exclusive scan of n ints in two launches; *total receives the sum.
  // nDev (or null): a count produced on the device, at most n.
  void scan(const int *in, int *out, int n, int *total, const int *nDev = nullptr) {
    if (n <= 0) {
      hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, stream, total, 1, 0);
      return;
    }
    const int nb = (n + 1023) / 1024;
    DBuf<int> &sums = scanTmp[0];
    if (sums.n < size_t(nb)) {
      // a replayed graph holds the buffer's address (the pack sizes it for
      // every scan a step makes, so a step being captured never grows it)
      if (graphCapturing) throw ApiError(AVBD_E_STATE, "scan buffer too small for a captured step");
      dropGraphs();
      sums.reserve(nb);
    }
    hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(256), 0, stream, in, out, n, sums.p, nDev);
    hipLaunchKernelGGL(k_scan_finish, dim3(nb), dim3(256), 0, stream, out, n, (const int *)sums.p, nb, total, nDev);
  }

  // Zero n ints: a memset, or in a step captured whole a fill kernel (memset
  // nodes are kept out of replayed graphs).
  void zero(int *p, long n, bool kernel) {
    if (n <= 0) return;
    if (kernel) hipLaunchKernelGGL(k_fill, dim3(blocks(n)), dim3(256), 0, stream, p, int(n), 0);
    else HIPCHECK(hipMemsetAsync(p, 0, sizeof(int) * size_t(n), stream));
  }
  // a grid for threads that stride over a count known only on the device
  static int strideBlocks(long n) { return std::max(1, std::min(blocks(n), 2048)); }

  void readScal(int first, int count) {
    HIPCHECK(hipMemcpyAsync(hscal + first, scal.p + first, sizeof(int) * count, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }

  void record(int phase) {
    if (profiling) HIPCHECK(hipEventRecord(evs[phase], stream));
  }

  // ------------------------------------------------------------ adjacency
  // Force lists (CSR) and per-body work lists. Their totals are bounded by
  // 2 (manifolds + joints) and 2 (8 manifolds + joints), for which the
  // buffers are sized at pack, so neither needs a host round trip.
  // dev: inside a step captured whole, the live count is read on the device
  // (scal[2]) and the grids stride over it.
  void buildAdjacency(bool dev = false) {
    const long nf = dev ? long(J) + manCap : long(J) + nLive;  // dev: an upper bound
    zero(deg.p, N, dev);
    const int4 *head = store[cur].head.p;
    const int *nLiveDev = dev ? scal.p + 2 : nullptr;
    const int gf = dev ? strideBlocks(nf) : blocks(nf), gr = dev ? strideBlocks(2 * nf) : blocks(2 * nf);
    if (nf > 0)
      hipLaunchKernelGGL(k_adj<0>, dim3(gf), dim3(256), 0, stream, nLive, nLiveDev, J, live[cur].p, head,
                         joints.p, (const BodyProp *)props.p, deg.p, adjCursor.p, adjKeyFill.p, adjOtherFill.p,
                         adjOwner.p);
    scan(deg.p, adjStart.p, N, adjStart.p + N);
    HIPCHECK(hipMemcpyAsync(adjCursor.p, adjStart.p, sizeof(int) * N, hipMemcpyDeviceToDevice, stream));
    if (nf > 0) {
      hipLaunchKernelGGL(k_adj<1>, dim3(gf), dim3(256), 0, stream, nLive, nLiveDev, J, live[cur].p, head,
                         joints.p, (const BodyProp *)props.p, deg.p, adjCursor.p, adjKeyFill.p, adjOtherFill.p,
                         adjOwner.p);
      // at most two entries per force
      hipLaunchKernelGGL(k_adj_rank, dim3(gr), dim3(256), 0, stream, int(2 * nf), N, props.p, adjStart.p,
                         (const int *)adjOwner.p, (const int *)adjKeyFill.p, (const int *)adjOtherFill.p, adjKey.p,
                         adjOther.p);
    }
  }

  void buildItems() {
    const int4 *head = store[cur].head.p;
    const int cap = int(items.n);
    hipLaunchKernelGGL(k_items<0>, dim3(blocks(N)), dim3(256), 0, stream, N, J, props.p, adjStart.p, adjKey.p,
                       live[cur].p, head, joints.p, itemCnt.p, itemStart.p, items.p, cap, scal.p + 5, itemJoints.p);
    scan(itemCnt.p, itemStart.p, N, itemStart.p + N);
    hipLaunchKernelGGL(k_items<1>, dim3(blocks(N)), dim3(256), 0, stream, N, J, props.p, adjStart.p, adjKey.p,
                       live[cur].p, head, joints.p, itemCnt.p, itemStart.p, items.p, cap, scal.p + 5, itemJoints.p);
  }

  // ------------------------------------------------------------ colouring
  // cmeta: [0..2] change counters, [3] rounds, [4] colours, [8..8+257) bucket
  // starts. Relaxation rounds go out in batches (colorBatch round + check
  // pairs) and reach the greedy fixed point; rounds after convergence return
  // at once. The bucket kernels follow the batch, so when it converged a single
  // readback brings the colour count and buckets to the host. The batch is
  // sized from the previous step's round count.
  // graphSame: the force graph is the one the colour array was computed for
  // (the step kept every manifold and made none, and there are no joints that
  // could have broken): one round then confirms the fixed point instead of a
  // batch re-deriving it. Pairs the broad phase offers that make no contact
  // do not change the graph.
  // The colour buckets of the colour array (colour count in cmeta[4], starts
  // from cmeta[8]), the sweep's work entries and the first-visitor flags.
  void bucketize(bool dev = false) {
    zero(cmeta.p + 4, 1, dev);
    const int nb = blocks(N);
    hipLaunchKernelGGL(k_color_hist, dim3(nb), dim3(256), 0, stream, N, color[0].p, chist.p, cmeta.p + 4);
    hipLaunchKernelGGL(k_color_offsets, dim3(MAX_COLORS), dim3(256), 0, stream, nb, chist.p, coffs.p, ctotal.p);
    hipLaunchKernelGGL(k_color_starts, dim3(1), dim3(MAX_COLORS), 0, stream, ctotal.p, cmeta.p + 8,
                       cmeta.p + CMETA_HEAVY, (const int *)(itemStart.p + N), nDynamic, heavyMin());
    hipLaunchKernelGGL(k_color_scatter, dim3(nb), dim3(256), 0, stream, N, color[0].p, coffs.p, cmeta.p + 8,
                       cbodies.p, itemStart.p, itemJoints.p, cwork.p, cmeta.p + CMETA_HEAVY, hwork.p);
    hipLaunchKernelGGL(k_item_flags, dim3(blocks(long(N) * 8)), dim3(256), 0, stream, N, color[0].p, itemStart.p,
                       items.p);
  }

  // The dynamic bodies sorted by the claim order of claimsBefore (index;
  // lowbias32 of the index within the scene, ties to the lower index; a
  // caller's rank, ties to the lower index), for k_color_dag.
  void ensureClaimOrder(int order) {
    if (corderFor == order && corder.p) return;
    std::vector<int> ids;
    ids.reserve(size_t(N));
    for (int i = 0; i < N; i++)
      if (hprops[size_t(i)].mass > 0) ids.push_back(i);
    if (order != AVBD_ORDER_INDEX) {
      std::vector<int64_t> key(size_t(N), 0);
      for (int i : ids) {
        if (order == AVBD_ORDER_CUSTOM) {
          key[size_t(i)] = hrank[size_t(i)];
        } else {
          uint32_t x = uint32_t(i - hsceneStart[size_t(hprops[size_t(i)].scene)]);
          x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
          key[size_t(i)] = int64_t(x);
        }
      }
      std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return key[size_t(a)] < key[size_t(b)]; });
    }
    corder.reserve(std::max<size_t>(ids.size(), 1));
    if (!ids.empty())
      HIPCHECK(hipMemcpy(corder.p, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice));
    nDynOrder = int(ids.size());
    corderFor = order;
  }
  // The colouring in one launch: state reset, the claim-order pass and the
  // report (cmeta[1]: every body coloured; cmeta[3]: rounds); no host read.
  void colorDag(int order) {
    dagState.reserve(size_t(N));
    hipLaunchKernelGGL(k_color_dag_init, dim3(blocks(N)), dim3(256), 0, stream, N, props.p, dagState.p, cmeta.p);
    const int waves = std::max(1, (nDynOrder + 63) / 64);
    if (nDynOrder > 0)
      hipLaunchKernelGGL(k_color_dag, dim3(blocks(long(waves) * 64)), dim3(256), 0, stream, nDynOrder,
                         (const int *)corder.p, (const BodyProp *)props.p, (const int *)adjStart.p,
                         (const int *)adjOther.p, color[0].p, dagState.p, order, (const int *)sceneStart.p,
                         (const int *)claimRank.p, cmeta.p, scal.p + 5, 2 * N + 4096);
    hipLaunchKernelGGL(k_color_dag_close, dim3(1), dim3(1), 0, stream, nDynOrder, cmeta.p);
  }
  bool dagUsable(bool uploaded, bool confirm) const {
    return dagColoring && !uploaded && !confirm && roundCap <= 0 && colorCap <= 0;
  }

  void colorize(bool graphSame) {
    HIPCHECK(hipMemsetAsync(cmeta.p, 0, sizeof(int) * 8, stream));
    const int maxRounds = N + 2;
    const int order = useRank ? AVBD_ORDER_CUSTOM : prm.color_order;
    bool uploaded = useUpload;
    const bool confirm = graphSame && !uploaded && colorFor == order;
    // the colouring in one launch; the relaxation rounds below take over if
    // it reports a body it could not colour
    bool dagDone = false;
    if (dagUsable(uploaded, confirm)) {
      ensureClaimOrder(order);
      colorDag(order);
      bucketize();
      HIPCHECK(hipMemcpyAsync(hmeta.data(), cmeta.p, sizeof(int) * hmeta.size(), hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      dagDone = hmeta[1] != 0;
      if (!dagDone) {
        dagColoring = false;
        HIPCHECK(hipMemsetAsync(cmeta.p, 0, sizeof(int) * 8, stream));
      }
    }
    for (int launched = 0; !dagDone;) {
      if (uploaded) {
        // the uploaded colouring, if it is proper for this step's forces
        hipLaunchKernelGGL(k_color_install, dim3(blocks(N)), dim3(256), 0, stream, N, props.p, uploadColor.p,
                           color[0].p, cmeta.p);
        hipLaunchKernelGGL(k_color_validate, dim3(blocks(N)), dim3(256), 0, stream, N, props.p, adjStart.p,
                           adjOther.p, color[0].p, cmeta.p + 5, scal.p + 5);
      } else {
        int batch = confirm && launched == 0 ? 1 : colorBatch;
        if (roundCap > 0) batch = std::max(0, std::min(batch, roundCap - launched));
        for (int r = 0; r < batch; r++) {
          hipLaunchKernelGGL(k_color_round, dim3(blocks(N)), dim3(256), 0, stream, N, props.p, adjStart.p,
                             adjOther.p, color[0].p, cmeta.p, order, sceneStart.p, scal.p + 5, claimRank.p);
        }
        launched += batch;
      }
      const bool capped = !uploaded && roundCap > 0 && launched >= roundCap;
      if (colorCap > 0)
        hipLaunchKernelGGL(k_color_cap, dim3(blocks(N)), dim3(256), 0, stream, N, color[0].p, colorCap, scal.p + 5,
                           (int *)nullptr);
      bucketize();
      HIPCHECK(hipMemcpyAsync(hmeta.data(), cmeta.p, sizeof(int) * hmeta.size(), hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      if (uploaded && hmeta[5]) {
        // not proper for this step (reported in the overflow bits): colour greedily
        uploaded = false;
        HIPCHECK(hipMemsetAsync(cmeta.p, 0, sizeof(int) * 4, stream));
        continue;
      }
      if (hmeta[1]) break;
      if (capped) {
        // the round limit ends the relaxation where it stands: a colouring
        // that may not be proper (reported in bit 32 of the overflow word)
        hipLaunchKernelGGL(k_or_flag, dim3(1), dim3(1), 0, stream, scal.p + 5, 32);
        break;
      }
      if (launched > maxRounds) throw ApiError(AVBD_E_STATE, "colouring did not converge");
    }
    colorRounds = hmeta[3];
    if (!dagDone && !uploaded && !confirm && roundCap <= 0) colorBatch = std::min(64, std::max(4, colorRounds + 2));
    const bool limited = (roundCap > 0 && !hmeta[1]) || colorCap > 0;
    colorFor = uploaded || limited ? -1 : order;
    numColors = hmeta[4];
    hcstart.assign(hmeta.begin() + 8, hmeta.begin() + 8 + numColors + 1);
    hheavy.assign(hmeta.begin() + CMETA_HEAVY, hmeta.begin() + CMETA_HEAVY + MAX_COLORS);
  }

  // The colouring inside a step captured whole: R relaxation rounds (those
  // after the fixed point return at once), the cap of E colours, the check
  // that reports a miss, and the buckets; no host read.
  void colorizeDev(int E, int R, bool ownRounds, bool ownCap) {
    zero(cmeta.p, 8, true);
    const int order = useRank ? AVBD_ORDER_CUSTOM : prm.color_order;
    if (dagUsable(useUpload, false) && corderFor == order) {
      // the colouring in one launch (its claim order built before capture)
      colorDag(order);
    } else {
      for (int r = 0; r < R; r++)
        hipLaunchKernelGGL(k_color_round, dim3(blocks(N)), dim3(256), 0, stream, N, props.p, adjStart.p, adjOther.p,
                           color[0].p, cmeta.p, order, sceneStart.p, scal.p + 5, claimRank.p);
    }
    hipLaunchKernelGGL(k_color_cap, dim3(blocks(N)), dim3(256), 0, stream, N, color[0].p, E, scal.p + 5,
                       pc.p + PC_SEEN);
    hipLaunchKernelGGL(k_pipe_color_check, dim3(1), dim3(1), 0, stream, (const int *)cmeta.p, pc.p, scal.p + 5, E,
                       ownRounds ? 1 : 0, ownCap ? 1 : 0);
    bucketize(true);
  }


  // bodies of colour c at the last colouring the host read (an estimate for
  // a graph's launch shapes; 0 for a colour it did not see)
  long colorCount(int c) const {
    if (pipeCapturing) return c >= 0 && c < int(pipeEst.size()) ? pipeEst[c] : 0;
    if (c < 0 || c + 1 >= int(hcstart.size())) return 0;
    return hcstart[c + 1] - hcstart[c];
  }

  // The solve loop's kernels take their counts from device memory and stride,
  // so one captured graph serves every step with that colour count; the
  // launch shapes are chosen from the colour sizes at capture.
  // A team per body of colour c, so that no team sweeps two bodies one after
  // the other (colours differ in size: the pyramid's first colours hold 1.7x
  // the average). A captured graph is replayed on later steps whose colours
  // may have grown: it gets a quarter more teams (and teams stride over any
  // excess, so a larger colour is still swept whole).
  int primalBlocks(int c, int tm, int blk = 256) const {
    const long cnt = std::max(colorCount(c), 1L);
    const long teams = graphCapturing ? cnt + cnt / 4 + 64 : cnt;
    return std::max(1, std::min(blocks(teams * tm, blk), 8192 * (256 / blk)));
  }
  // heavy bodies of colour c at the last colouring the host read
  long heavyCount(int c) const {
    if (pipeCapturing) return c >= 0 && c < int(pipeEstH.size()) ? pipeEstH[c] : 0;
    return c >= 0 && c < int(hheavy.size()) ? hheavy[c] : 0;
  }
  // blocks of wide teams (tw lanes) for colour c's heavy bodies: at least
  // one, which finds the count on the device and strides over it
  int heavyBlocks(int c, int tw, int blk) const {
    const long cnt = heavyCount(c);
    const long teams = graphCapturing ? cnt + cnt / 4 + 8 : cnt;
    return std::max(1, std::min(blocks(teams * tw, blk), 1024 * (256 / blk)));
  }
  // the least item count that makes a body heavy (the device raises it to
  // twice the mean): twice the team, so that a body swept by the usual team
  // holds at most two items per lane; 0 (no split) for the exact precision
  // and teams too narrow or too wide to split
  int heavyMin() const {
    if (!heavySplit || prec == AVBD_PREC_EXACT || team < 4 || team * HEAVY_WIDEN > 64) return 0;
    return 2 * team;
  }
  int dualBlocks() const { return std::max(1, std::min(blocks(manCap, 128), 2048)); }

  template <class R, bool EXACT, int TEAM>
  void launchPrimal(int c, const SolveArgs &sa, ManBuf<R> mb, const StepParams &P, double alpha, int dual,
                    double dualAlpha) {
    if (!graphCapturing && colorCount(c) <= 0) return;
    const double bytes = profiling ? double(hcbytes[c]) + (dual ? double(hcbytes[MAX_COLORS + c]) : 0.0) : 0.0;
    // the tuned workgroup; the one-lane exact sweep keeps 256
    const int blk = EXACT ? 256 : sweepBlock;
    // the team modes' heavy bodies go to blocks of their own, first
    constexpr bool split = heavySplitTeam<R, EXACT, TEAM>();
    const int light = primalBlocks(c, TEAM, blk), wide = split ? heavyBlocks(c, HEAVY_WIDEN * TEAM, blk) : 0;
    timedLaunch(AVBD_K_PRIMAL, bytes, gl() ? k_primal<R, EXACT, TEAM, true> : k_primal<R, EXACT, TEAM, false>,
                dim3(light + wide), dim3(blk), c, (const int *)(cmeta.p + 8), (const int4 *)cwork.p,
                (const int *)(cmeta.p + CMETA_HEAVY), (const int4 *)hwork.p, wide, sa, mb, (R *)DQ.p, P, alpha,
                dual, dualAlpha);
  }

  // Team shape for colour c: a colour too small to give every SIMD two waves
  // at the tuned shape gets wider teams (fewer contacts per lane, so a shorter
  // chain of gathers per body); the smallest colours, one launch each like
  // the large ones, are where the sweep waits longest per body. Only while
  // each lane of the wider team keeps at least one work item on average (a
  // contact or joint is one item for each dynamic body it touches): beyond
  // that a wider team only idles its lanes and lengthens its reduction.
  int teamFor(int c) const {
    int tm = team;
    if (tm < 4) return tm;
    const long cnt = colorCount(c);
    const double perBody = double(2L * contactsTotal + 2L * J) / double(std::max(nDynamic, 1));
    while (tm < 64 && cnt * long(tm) * 2 <= 2L * simds * 64 && tm * 2 <= perBody) tm *= 2;
    return tm;
  }

  // One colour of the sweep; dual != 0 folds in the previous iteration's
  // contact dual update (evaluated with dualAlpha).
  template <class R, bool EXACT>
  void primal(int c, const SolveArgs &sa, ManBuf<R> mb, const StepParams &P, double alpha, int tm, int dual,
              double dualAlpha) {
    if constexpr (EXACT) {
      if (!exactTeam) { launchPrimal<R, true, 1>(c, sa, mb, P, alpha, dual, dualAlpha); return; }
      if (!graphCapturing && colorCount(c) <= 0) return;
      // a team of lanes per body, the additions kept in the reference's
      // order; bodies with joints then take the one-lane sweep
      const double bytes = profiling ? double(hcbytes[c]) + (dual ? double(hcbytes[MAX_COLORS + c]) : 0.0) : 0.0;
      const int blk = exactBlock;
      auto xk = blk == 64    ? (gl() ? k_primal_xt<true, 64> : k_primal_xt<false, 64>)
                : blk == 128 ? (gl() ? k_primal_xt<true, 128> : k_primal_xt<false, 128>)
                             : (gl() ? k_primal_xt<true, 256> : k_primal_xt<false, 256>);
      timedLaunch(AVBD_K_PRIMAL, bytes, xk, dim3(primalBlocks(c, XT_TEAM, blk)), dim3(blk), c,
                  (const int *)(cmeta.p + 8), (const int4 *)cwork.p, sa, mb, (double *)DQ.p, P, alpha, dual,
                  dualAlpha);
      if (J > 0) {
        SolveArgs sj = sa;
        sj.jointBodiesOnly = 1;
        const int grid = primalBlocks(c, 1);
        timedLaunch(AVBD_K_PRIMAL, 0.0, gl() ? k_primal<double, true, 1, true> : k_primal<double, true, 1, false>,
                    dim3(grid), dim3(256), c, (const int *)(cmeta.p + 8), (const int4 *)cwork.p,
                    (const int *)(cmeta.p + CMETA_HEAVY), (const int4 *)hwork.p, 0, sj, mb, (double *)DQ.p, P,
                    alpha, dual, dualAlpha);
      }
      return;
    }
    switch (tm) {
      case 1: launchPrimal<R, false, 1>(c, sa, mb, P, alpha, dual, dualAlpha); break;
      case 4: launchPrimal<R, false, 4>(c, sa, mb, P, alpha, dual, dualAlpha); break;
      case 8: launchPrimal<R, false, 8>(c, sa, mb, P, alpha, dual, dualAlpha); break;
      case 32: launchPrimal<R, false, 32>(c, sa, mb, P, alpha, dual, dualAlpha); break;
      case 64: launchPrimal<R, false, 64>(c, sa, mb, P, alpha, dual, dualAlpha); break;
      default: launchPrimal<R, false, 16>(c, sa, mb, P, alpha, dual, dualAlpha); break;
    }
  }

  // body initialisation, then the iterations (solver.js:366-489)
  template <class R, bool EXACT>
  void enqueueSolve(const SolveArgs &sa, ManBuf<R> mb, const StepParams &P, bool withInit) {
    if (withInit)
      hipLaunchKernelGGL((k_body_init<R>), dim3(blocks(N)), dim3(256), 0, stream, N, props.p, X.p, X0.p, Y.p, V.p,
                         (R *)DQ.p, P);
    if (rec) recordSnapshot(0, 0);
    // The contact dual update of iteration it runs inside the primal sweep of
    // iteration it + 1 (first visitors apply it); after the last iteration it
    // runs as its own pass, which also records C for the next step.
    const int total = prm.iterations + (prm.post_stabilize ? 1 : 0);
    auto alphaOf = [&](int it) { return prm.post_stabilize ? (it < prm.iterations ? 1.0 : 0.0) : prm.alpha; };
    for (int it = 0; it < total; it++) {
      const double alpha = alphaOf(it);
      const int fused = it >= 1 && it < prm.iterations ? 1 : 0;
      const double dualAlpha = fused ? alphaOf(it - 1) : alpha;
      for (int c = 0; c < numColors; c++) primal<R, EXACT>(c, sa, mb, P, alpha, teamFor(c), fused, dualAlpha);
      // recording: the poses after this sweep, and the contacts' k, lambda as
      // this sweep used them (the folded update of the previous iteration has
      // reached every contact once this sweep is done)
      if (rec && it < prm.iterations) recordSnapshot(it + 1, fused ? it : -1);
      if (it < prm.iterations) {
        const bool last = it == prm.iterations - 1;
        if (last && (graphCapturing || nLive > 0))
          timedLaunch(AVBD_K_DUAL, profiling ? dualBytes() : 0.0, gl() ? k_dual<R, EXACT, true> : k_dual<R, EXACT, false>,
                      dim3(dualBlocks()), dim3(128), (const int *)(scal.p + 2), (const int *)live[cur].p, mb,
                      (const R *)DQ.p, P, alpha, last ? 1 : 0);
        if (J > 0) hipLaunchKernelGGL(k_joint_dual, dim3(blocks(J)), dim3(256), 0, stream, J, joints.p, X.p, P, alpha);
      }
      if (rec && it == prm.iterations - 1) recordSnapshot(-1, it + 1);
      if (it == prm.iterations - 1)
        hipLaunchKernelGGL(k_velocity, dim3(blocks(N)), dim3(256), 0, stream, N, props.p, X.p, X0.p, V.p, prm.dt);
    }
  }

  // The batch's solve in one launch (k_solve_scenes), then the contact dual
  // of the last iteration (k_dual, which also records C). Used when the
  // packed batch qualifies, post-stabilisation is off and no adjoint step is
  // being recorded.
  bool useScenes() const { return sceneSolveOK && sceneSolveCfg >= 0 && !prm.post_stabilize && !rec; }

  // Lanes per body in the one-launch batch solve: the tuned shape, else 4
  // (the batch's scenes hold a few contacts or joints per body: the 256
  // scenes of chains and stacks solve in 0.61 ms at 4, 0.97 at 8, 2.1 at 16)
  int sceneTeamShape() const {
    const int t = sceneTeam > 0 ? sceneTeam : 4;
    return t <= 4 ? 4 : t <= 8 ? 8 : 16;
  }
  // Workgroup of the one-launch batch solve (lanes per scene): the tuned
  // size, else 256. With TEAM lanes per body, a workgroup updates
  // BLOCK / TEAM bodies of a colour at once (the colour tile); a scene whose
  // colour holds more takes several passes.
  int sceneBlockShape() const { return sceneBlock == 128 || sceneBlock == 512 ? sceneBlock : 256; }

  template <class R, int BLOCK>
  static auto sceneKernel(int tm, bool law) {
    auto k = law ? k_solve_scenes<R, false, 16, true, BLOCK> : k_solve_scenes<R, false, 16, false, BLOCK>;
    if (tm == 4) k = law ? k_solve_scenes<R, false, 4, true, BLOCK> : k_solve_scenes<R, false, 4, false, BLOCK>;
    if (tm == 8) k = law ? k_solve_scenes<R, false, 8, true, BLOCK> : k_solve_scenes<R, false, 8, false, BLOCK>;
    return k;
  }

  template <class R, bool EXACT>
  void launchScenes(const SolveArgs &sa, ManBuf<R> mb, const StepParams &P, int tm, int blk) {
    SceneSolve Q;
    Q.sceneStart = sceneStart.p; Q.sceneJointStart = sceneJointStart.p; Q.cmeta = cmeta.p; Q.cwork = cwork.p;
    Q.V = V.p; Q.joints = joints.p; Q.iterations = prm.iterations; Q.alpha = prm.alpha;
    // algorithmic bytes: every sweep of the step (the folded dual update in
    // all but the first iteration)
    double bytes = 0.0;
    if (profiling)
      for (int c = 0; c < numColors; c++)
        bytes += double(prm.iterations) * double(hcbytes[c]) + double(prm.iterations - 1) * double(hcbytes[MAX_COLORS + c]);
    R *dq = (R *)DQ.p;
    if constexpr (EXACT) {
      // the exact team's LDS rows are sized for 256 lanes
      timedLaunch(AVBD_K_PRIMAL, bytes,
                  gl() ? k_solve_scenes<R, true, 1, true, 256> : k_solve_scenes<R, true, 1, false, 256>, dim3(S),
                  dim3(256), Q, sa, mb, dq, P);
    } else {
      const auto k = blk == 128 ? sceneKernel<R, 128>(tm, gl())
                     : blk == 512 ? sceneKernel<R, 512>(tm, gl())
                                  : sceneKernel<R, 256>(tm, gl());
      timedLaunch(AVBD_K_PRIMAL, bytes, k, dim3(S), dim3(blk == 128 || blk == 512 ? blk : 256), Q, sa, mb, dq, P);
    }
  }

  template <class R, bool EXACT>
  void enqueueScenes(const SolveArgs &sa, ManBuf<R> mb, const StepParams &P) {
    launchScenes<R, EXACT>(sa, mb, P, sceneTeamShape(), sceneBlockShape());
    if (graphCapturing || nLive > 0)
      timedLaunch(AVBD_K_DUAL, profiling ? dualBytes() : 0.0, gl() ? k_dual<R, EXACT, true> : k_dual<R, EXACT, false>,
                  dim3(dualBlocks()), dim3(128), (const int *)(scal.p + 2), (const int *)live[cur].p, mb,
                  (const R *)DQ.p, P, prm.alpha, 1);
  }

  // The autotuner for the one-launch batch solve: the whole solve of this
  // step timed for each team shape (lanes per body: 4, 8, 16) and workgroup
  // (lanes per scene: 128, 256, 512; the tile of bodies a workgroup updates
  // at once is their ratio), restoring everything it writes between trials;
  // the fastest pair is kept.
  template <class R>
  void autotuneScenes(const SolveArgs &sa, ManBuf<R> mb, const StepParams &P) {
    static const int cands[3] = {4, 8, 16};
    static const int blks[3] = {128, 256, 512};
    const size_t xb = sizeof(double) * size_t(N) * XS, vb = sizeof(double) * size_t(N) * VS,
                 qb = size_t(N) * 8 * rsize(), jb = sizeof(JointDev) * size_t(J);
    const size_t nT = size_t(std::max(lastNTot, 1)), hb = nT * MAXC * HOT * rsize(), sb = nT * MAXC * sizeof(int32_t);
    ManStore &m = store[cur];
    DBuf<char> bX, bX0, bY, bV, bQ, bJ, bH, bS;
    bX.reserve(xb); bX0.reserve(xb); bY.reserve(xb); bV.reserve(vb); bQ.reserve(qb); bJ.reserve(std::max<size_t>(jb, 1));
    bH.reserve(hb); bS.reserve(sb);
    auto copy = [&](void *d, const void *src, size_t n) {
      if (n) HIPCHECK(hipMemcpyAsync(d, src, n, hipMemcpyDeviceToDevice, stream));
    };
    auto save = [&]() {
      copy(bX.p, X.p, xb); copy(bX0.p, X0.p, xb); copy(bY.p, Y.p, xb); copy(bV.p, V.p, vb); copy(bQ.p, DQ.p, qb);
      copy(bJ.p, joints.p, jb); copy(bH.p, m.hot.p, hb); copy(bS.p, m.stick.p, sb);
    };
    auto restore = [&]() {
      copy(X.p, bX.p, xb); copy(X0.p, bX0.p, xb); copy(Y.p, bY.p, xb); copy(V.p, bV.p, vb); copy(DQ.p, bQ.p, qb);
      copy(joints.p, bJ.p, jb); copy(m.hot.p, bH.p, hb); copy(m.stick.p, bS.p, sb);
    };
    save();
    hipEvent_t a = takeEvent(), b = takeEvent();
    const bool prof = profiling;
    profiling = false;
    int best = sceneTeamShape(), bestBlk = sceneBlockShape();
    float bestMs = 1e30f;
    for (int k = 0; k < 6; k++) tuneMs[k] = 0.0f;
    std::memset(tuneGrid, 0, sizeof(tuneGrid));
    for (int k = 0; k < 3; k++)
      for (int g = 0; g < 3; g++) {
        HIPCHECK(hipEventRecord(a, stream));
        launchScenes<R, false>(sa, mb, P, cands[k], blks[g]);
        HIPCHECK(hipEventRecord(b, stream));
        HIPCHECK(hipEventSynchronize(b));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, a, b));
        tuneGrid[k + 1][g + 1] = ms;
        if (tuneMs[k + 1] == 0.0f || ms < tuneMs[k + 1]) tuneMs[k + 1] = ms;
        if (ms < bestMs) { bestMs = ms; best = cands[k]; bestBlk = blks[g]; }
        restore();
      }
    HIPCHECK(hipStreamSynchronize(stream));
    bX.release(); bX0.release(); bY.release(); bV.release(); bQ.release(); bJ.release(); bH.release(); bS.release();
    profiling = prof;
    evPool.push_back(a);
    evPool.push_back(b);
    if (best != sceneTeam || bestBlk != sceneBlock) dropGraphs();
    sceneTeam = best;
    sceneBlock = bestBlk;
  }

  // Graph executables are destroyed only when no replay of them is pending:
  // a replay still queued or running on the stream, or a tool's completion
  // callback for its kernels, must not outlive the executable it came from.
  void quiesceForGraphs() {
    if (!stream || graphs.empty()) return;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &st) == hipSuccess && st == hipStreamCaptureStatusNone)
      (void)hipStreamSynchronize(stream);
  }
  void dropGraphs() {
    quiesceForGraphs();
    for (auto &g : graphs) {
      (void)hipGraphExecDestroy(g.second.second);
      (void)hipGraphDestroy(g.second.first);
    }
    graphs.clear();
  }

  // Replay (capturing on first use) the solve loop for this colour count and
  // manifold buffer parity.
  template <class R, bool EXACT>
  void solveGraph(const SolveArgs &sa, ManBuf<R> mb, const StepParams &P) {
    const long long key = (long long)numColors * 4 + cur;
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      hipGraph_t g = nullptr;
      HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      graphCapturing = true;
      try {
        enqueueSolve<R, EXACT>(sa, mb, P, true);
      } catch (...) {
        graphCapturing = false;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      graphCapturing = false;
      HIPCHECK(hipStreamEndCapture(stream, &g));
      hipGraphExec_t ex = nullptr;
      const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      if (e != hipSuccess) (void)hipGraphDestroy(g);
      HIPCHECK(e);
      it = graphs.emplace(key, std::make_pair(g, ex)).first;
    }
    HIPCHECK(hipGraphLaunch(it->second.second, stream));
  }

  // The one-launch batch solve as a replayed graph: its launches take every
  // count from device memory, so one graph per manifold buffer parity serves
  // every step (keys above 4 * (MAX_COLORS + 1) are these).
  template <class R, bool EXACT>
  void sceneGraph(const SolveArgs &sa, ManBuf<R> mb, const StepParams &P) {
    const long long key = 4LL * (MAX_COLORS + 1) + cur;
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      hipGraph_t g = nullptr;
      HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      graphCapturing = true;
      try {
        enqueueScenes<R, EXACT>(sa, mb, P);
      } catch (...) {
        graphCapturing = false;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      graphCapturing = false;
      HIPCHECK(hipStreamEndCapture(stream, &g));
      hipGraphExec_t ex = nullptr;
      const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      if (e != hipSuccess) (void)hipGraphDestroy(g);
      HIPCHECK(e);
      it = graphs.emplace(key, std::make_pair(g, ex)).first;
    }
    HIPCHECK(hipGraphLaunch(it->second.second, stream));
  }

  // One primal sweep (every colour, first iteration's alpha) per team shape,
  // timed with events; the body state it changes (X, DQ) is restored after
  // each trial. Keeps the fastest shape.
  // Workgroups of 64, 128 and 256 lanes are tried with each team shape (a
  // colour's teams are spread over the CUs in workgroup-sized pieces).
  // each colour's heavy-body count from the last bucketize, for the launch
  // shapes of direct launches
  void readHeavy() {
    hheavy.resize(MAX_COLORS);
    HIPCHECK(hipMemcpyAsync(hheavy.data(), cmeta.p + CMETA_HEAVY, sizeof(int) * MAX_COLORS, hipMemcpyDeviceToHost,
                            stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }

  template <class R>
  void autotuneT(const SolveArgs &sa, ManBuf<R> mb, const StepParams &P) {
    static const int cands[6] = {1, 4, 8, 16, 32, 64};
    static const int blks[3] = {64, 128, 256};
    const size_t xb = sizeof(double) * size_t(N) * XS, qb = size_t(N) * 8 * rsize();
    Xbk.reserve(xb);
    DQbk.reserve(qb);
    HIPCHECK(hipMemcpyAsync(Xbk.p, X.p, xb, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(DQbk.p, DQ.p, qb, hipMemcpyDeviceToDevice, stream));
    hipEvent_t a = takeEvent(), b = takeEvent();
    const bool prof = profiling;
    profiling = false;
    int best = team, bestBlk = sweepBlock;
    float bestMs = 1e30f;
    const int keepBlk = sweepBlock;
    std::memset(tuneGrid, 0, sizeof(tuneGrid));
    const int keepTeam = team;
    for (int k = 0; k < 6; k++) {
      tuneMs[k] = 0.0f;
      // the heavy-body lists this team would sweep with
      team = cands[k];
      bucketize();
      readHeavy();
      for (int g = 0; g < 3; g++) {
        sweepBlock = blks[g];
        HIPCHECK(hipEventRecord(a, stream));
        for (int c = 0; c < numColors; c++) primal<R, false>(c, sa, mb, P, prm.alpha, cands[k], 0, prm.alpha);
        HIPCHECK(hipEventRecord(b, stream));
        HIPCHECK(hipEventSynchronize(b));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, a, b));
        tuneGrid[k][g] = ms;
        if (tuneMs[k] == 0.0f || ms < tuneMs[k]) tuneMs[k] = ms;
        if (ms < bestMs) { bestMs = ms; best = cands[k]; bestBlk = blks[g]; }
        HIPCHECK(hipMemcpyAsync(X.p, Xbk.p, xb, hipMemcpyDeviceToDevice, stream));
        HIPCHECK(hipMemcpyAsync(DQ.p, DQbk.p, qb, hipMemcpyDeviceToDevice, stream));
      }
    }
    sweepBlock = keepBlk;
    team = keepTeam;
    profiling = prof;
    evPool.push_back(a);
    evPool.push_back(b);
    // the step's solve then runs with the chosen team's lists
    std::swap(team, best);
    bucketize();
    readHeavy();
    std::swap(team, best);
    if (best != team || bestBlk != sweepBlock) {
      dropGraphs();
      // the heavy-body threshold follows the team: the next step reads the
      // colours' heavy-body counts on the host, which the step graphs
      // captured after it are shaped for
      pipeReady = false;
    }
    team = best;
    sweepBlock = bestBlk;
  }

  // ---------------------------------------------------------------- step
  // 1. broad phase: the pair lists and their counts (scal[0], scal[1]).
  // dev: inside a step captured whole (fill kernels for the clears).
  void broadPhase(bool dev) {
    zero(bucketCnt.p, long(tableMask) + 1, dev);

    hipLaunchKernelGGL(k_bp_cells, dim3(blocks(N)), dim3(256), 0, stream, N, props.p, X.p, invh, tableMask, cellOf.p,
                       bucketCnt.p);
    scan(bucketCnt.p, bucketStart.p, int(tableMask + 1), scal.p + 7);
    HIPCHECK(hipMemcpyAsync(bucketCursor.p, bucketStart.p, sizeof(int) * (tableMask + 1), hipMemcpyDeviceToDevice,
                            stream));
    hipLaunchKernelGGL(k_bp_scatter, dim3(blocks(N)), dim3(256), 0, stream, N, cellOf.p, tableMask, bucketCursor.p,
                       entries.p, entryCell.p);
    BpArgs ba;
    ba.n = N; ba.J = J; ba.props = props.p; ba.X = X.p; ba.cellOf = cellOf.p; ba.bucketStart = bucketStart.p;
    ba.bucketCnt = bucketCnt.p; ba.entries = entries.p; ba.mask = tableMask; ba.sceneLargeStart = sceneLargeStart.p;
    ba.sceneLarges = sceneLarges.p; ba.sceneStart = sceneStart.p; ba.adjStart = adjStart.p; ba.adjKey = adjKey.p;
    ba.adjOther = adjOther.p; ba.joints = joints.p; ba.countRev = countRev.p; ba.startRev = startRev.p;
    ba.countNewRev = countNewRev.p; ba.startNewRev = startNewRev.p;
    ba.pairs = pairs.p; ba.newPairs = newPairs.p; ba.pairCap = pairCap;
    ba.entryCell = entryCell.p; ba.scratch = bpScratch.p;
    ba.skin = prm.contact_model ? skinOff.p : nullptr;
    zero(countRev.p, N, dev);
    zero(countNewRev.p, N, dev);
    hipLaunchKernelGGL(k_bp_team, dim3(blocks(N, 256 / BP_TEAM)), dim3(256), 0, stream, ba);
    if (nLarge) hipLaunchKernelGGL(k_bp_large<0>, dim3(nLarge), dim3(256), 0, stream, ba, largeList.p);
    scan(countRev.p, startRev.p, N, scal.p + 0);
    scan(countNewRev.p, startNewRev.p, N, scal.p + 1);
    // the write pass keeps every write below pairCap; an excess is reported
    // before anything reads the lists
    hipLaunchKernelGGL(k_bp_emit, dim3(blocks(N)), dim3(256), 0, stream, ba);
    if (nLarge) hipLaunchKernelGGL(k_bp_large<1>, dim3(nLarge), dim3(256), 0, stream, ba, largeList.p);
  }

  template <class R, bool EXACT>
  void stepT() {
    const StepParams P = sp();
    record(0);
    broadPhase(false);
    readScal(0, 2);
    nPairs = hscal[0];
    nNew = hscal[1];
    if (nPairs > pairCap) {
      nPairs = 0;
      needPairs = std::max(needPairs, hscal[0]);
      throw ApiError(AVBD_E_CAPACITY, "broad-phase pairs " + std::to_string(hscal[0]) + " exceed capacity " +
                                          std::to_string(pairCap));
    }
    record(1);

    // 2. narrow phase and force initialisation
    const int nOld = nLive, nTot = nOld + nNew;
    if (nTot > manCap) {
      needMan = std::max(needMan, nTot);
      throw ApiError(AVBD_E_CAPACITY, "manifolds " + std::to_string(nTot) + " exceed capacity " +
                                          std::to_string(manCap));
    }
    ManBuf<R> ob = store[cur].view<R>(), nb = store[1 - cur].view<R>();
    HIPCHECK(hipMemsetAsync(scal.p + 8, 0, sizeof(int), stream));
    if (nTot > 0)
      timedLaunch(AVBD_K_NARROW, 0.0, gl() ? k_narrow<R, true> : k_narrow<R, false>, dim3(blocks(nTot, 64)), dim3(64),
                  nOld, nNew, (const int *)nullptr,
                  (const int *)live[cur].p, ob, nb, (const int2 *)newPairs.p, (const BodyProp *)props.p,
                  (const double *)X.p, P, keep.p, scal.p + 8, srcSlot.p, (const BodyMat *)mat.p,
                  (const double *)V.p);
    if (J > 0) hipLaunchKernelGGL(k_joint_init, dim3(blocks(J)), dim3(256), 0, stream, J, joints.p, X.p, P);
    scan(keep.p, keepPos.p, nTot, scal.p + 2);
    if (nTot > 0)
      hipLaunchKernelGGL(k_live_compact, dim3(blocks(nTot)), dim3(256), 0, stream, nTot, (const int *)nullptr,
                         keep.p, keepPos.p, live[1 - cur].p);
    // how many of last step's manifolds survived (the exclusive scan at nOld)
    if (nNew > 0 && nOld > 0)
      HIPCHECK(hipMemcpyAsync(scal.p + 9, keepPos.p + nOld, sizeof(int), hipMemcpyDeviceToDevice, stream));
    readScal(2, 8);
    nLive = hscal[2];
    const int keptOld = nNew > 0 ? (nOld > 0 ? hscal[9] : 0) : nLive;
    contactsTotal = hscal[8];
    cur = 1 - cur;
    lastNTot = nTot;
    if (rec) recordContacts(nTot);
    record(2);

    // 3. topology: force lists, colouring, work lists
    buildAdjacency();
    buildItems();
    record(3);
    // the same force graph as last step: every old manifold kept, none made
    colorize(keptOld == nOld && nLive == nOld && J == 0);
    if (profiling && numColors > 0) {
      HIPCHECK(hipMemsetAsync(cbytes.p, 0, sizeof(unsigned long long) * 2 * MAX_COLORS, stream));
      hipLaunchKernelGGL(k_color_bytes, dim3(blocks(N)), dim3(256), 0, stream, N, J, props.p, color[colorCur].p,
                         adjStart.p, adjKey.p, itemStart.p, items.p, primalModel(), cbytes.p);
      hcbytes.resize(2 * MAX_COLORS);
      HIPCHECK(hipMemcpyAsync(hcbytes.data(), cbytes.p, sizeof(unsigned long long) * 2 * MAX_COLORS,
                              hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
    }
    record(4);
    if (rec) recordTopology();

    // 4. body initialisation and the iterations
    ManBuf<R> mb = store[cur].view<R>();
    SolveArgs sa;
    sa.props = props.p; sa.X = X.p; sa.X0 = X0.p; sa.Y = Y.p; sa.itemStart = itemStart.p; sa.items = items.p;
    sa.head = mb.head; sa.joints = joints.p; sa.nBodies = N; sa.jointBodiesOnly = 0;
    if (useScenes()) {
      if (tuneRequest && !EXACT) {
        autotuneScenes<R>(sa, mb, P);
        tuneRequest = false;
      }
      if (graphsCfg >= 0 && !profiling) sceneGraph<R, EXACT>(sa, mb, P);
      else enqueueScenes<R, EXACT>(sa, mb, P);
    } else if (tuneRequest && !EXACT) {
      hipLaunchKernelGGL((k_body_init<R>), dim3(blocks(N)), dim3(256), 0, stream, N, props.p, X.p, X0.p, Y.p, V.p,
                         (R *)DQ.p, P);
      autotuneT<R>(sa, mb, P);
      tuneRequest = false;
      enqueueSolve<R, EXACT>(sa, mb, P, false);
    } else if (graphsCfg >= 0 && !profiling && !rec) {
      solveGraph<R, EXACT>(sa, mb, P);
    } else {
      enqueueSolve<R, EXACT>(sa, mb, P, true);
    }
    record(5);
    HIPCHECK(hipGetLastError());
    steps++;
    pipeReady = true;
    if (profiling) {
      HIPCHECK(hipEventSynchronize(evs[5]));
      float ms;
      for (int ph = 0; ph < 4; ph++) {
        HIPCHECK(hipEventElapsedTime(&ms, evs[ph], evs[ph + 1]));
        profMs[ph] += ms;
      }
      HIPCHECK(hipEventElapsedTime(&ms, evs[4], evs[5]));
      profMs[4] += ms;
      HIPCHECK(hipEventElapsedTime(&ms, evs[0], evs[5]));
      profMs[5] += ms;
      profSteps++;
      harvest();
    }
  }

  // ------------------------------------------------- whole-step graphs
  // The step of stepT with every count kept on the device: the narrow phase,
  // compaction and force lists read the counts k_pipe_counts derived, the
  // colouring launches a fixed number of rounds and E colour buckets, the
  // solve E colour launches per iteration (a colour with no body returns at
  // once). Launch shapes come from the last step the host read. Nothing is
  // executed here: this is captured, then replayed.
  template <class R, bool EXACT>
  void enqueueStepDevice(int E, int rounds, bool ownRounds, bool ownCap) {
    const StepParams P = sp();
    broadPhase(true);
    hipLaunchKernelGGL(k_pipe_counts, dim3(1), dim3(1), 0, stream, scal.p, pc.p, pairCap, manCap);
    ManBuf<R> ob = store[cur].view<R>(), nb = store[1 - cur].view<R>();
    const long est = narrowEstimate();
    auto narrow = gl() ? k_narrow<R, true> : k_narrow<R, false>;
    hipLaunchKernelGGL(narrow, dim3(blocks(est, 64)), dim3(64), 0, stream, 0, 0, (const int *)pc.p, (const int *)live[cur].p, ob, nb, (const int2 *)newPairs.p,
                       (const BodyProp *)props.p, (const double *)X.p, P, keep.p, scal.p + 8, srcSlot.p,
                       (const BodyMat *)mat.p, (const double *)V.p);
    if (J > 0) hipLaunchKernelGGL(k_joint_init, dim3(blocks(J)), dim3(256), 0, stream, J, joints.p, X.p, P);
    scan(keep.p, keepPos.p, manCap, scal.p + 2, pc.p + PC_TOT);
    hipLaunchKernelGGL(k_live_compact, dim3(strideBlocks(est)), dim3(256), 0, stream, 0, (const int *)(pc.p + PC_TOT),
                       keep.p, keepPos.p, live[1 - cur].p);
    cur = 1 - cur;
    buildAdjacency(true);
    buildItems();
    colorizeDev(E, rounds, ownRounds, ownCap);
    ManBuf<R> mb = store[cur].view<R>();
    SolveArgs sa;
    sa.props = props.p; sa.X = X.p; sa.X0 = X0.p; sa.Y = Y.p; sa.itemStart = itemStart.p; sa.items = items.p;
    sa.head = mb.head; sa.joints = joints.p; sa.nBodies = N; sa.jointBodiesOnly = 0;
    if (useScenes()) enqueueScenes<R, EXACT>(sa, mb, P);
    else enqueueSolve<R, EXACT>(sa, mb, P, true);
    hipLaunchKernelGGL(k_pipe_stats, dim3(1), dim3(64), 0, stream, (const int *)scal.p, (const int *)cmeta.p, pc.p,
                       pipeRingDev, int(PIPE_SLOTS));
  }

  // threads of the narrow phase and compaction in a step graph: a quarter
  // above the manifold count it is sized for (more stride)
  long narrowEstimate() const {
    return std::min<long>(manCap, std::max<long>(4096, pipeTotEst + pipeTotEst / 4));
  }
  // the shapes of the next step graphs from the counts the host last read
  void pipeShapesFromHost() {
    // never below what earlier steps' graphs were shaped for: colour sizes
    // move from step to step, and every shape that falls short has the step
    // graphs captured again
    pipeEst.resize(MAX_COLORS, 0);
    for (int c = 0; c < numColors; c++) pipeEst[c] = std::max<long>(pipeEst[c], hcstart[c + 1] - hcstart[c]);
    pipeEstH.resize(MAX_COLORS, 0);
    for (int c = 0; c < std::min(numColors, int(hheavy.size())); c++)
      pipeEstH[c] = std::max<long>(pipeEstH[c], hheavy[c]);
    pipeTotEst = std::max<long>(pipeTotEst, lastNTot);
  }
  void dropStepGraphs() {
    quiesceForGraphs();
    for (auto it = graphs.begin(); it != graphs.end();) {
      if (it->first >= (1LL << 40)) {
        (void)hipGraphExecDestroy(it->second.second);
        (void)hipGraphDestroy(it->second.first);
        it = graphs.erase(it);
      } else {
        ++it;
      }
    }
  }

  // colour launches and relaxation rounds of the next graph step: the
  // caller's limits (GpuBackend.encColors / colorRounds) where set, else one
  // colour and four rounds above what the last step the host read needed,
  // raised when a step reports it needed more
  int pipeColorCount() const {
    if (useScenes()) return MAX_COLORS;  // one launch reads every colour
    if (colorCap > 0) return colorCap;
    const int want = std::max(pipeColors, numColors + 1);
    return std::max(1, std::min(MAX_COLORS, want));
  }
  int pipeRoundCount() const {
    // a step graph holds one launch per round: a caller's limit above the
    // N + 2 rounds any colouring needs launches no more than those
    if (roundCap > 0) return std::min(roundCap, N + 2);
    return std::min(N + 2, std::max({8, pipeRounds, colorRounds + 4}));
  }

  // Read ring slot k (its step has completed): misses raise the next graphs'
  // counts; a colour or manifold count past what the graphs were sized for
  // has them captured again (their teams and threads would stride).
  void takeSlot(int64_t k) {
    const int *o = pipeRing + (k % PIPE_SLOTS) * PIPE_STRIDE;
    const int miss = o[8];
    if (miss & 4) {
      needPairs = std::max(needPairs, o[10]);
      needMan = std::max(needMan, o[11]);
      pipeCapacity = true;
    }
    if (miss & 1) {
      // relaxation rounds short of the fixed point, or the one-launch
      // colouring unfinished: more rounds, and rounds from now on
      pipeRounds = std::min(N + 2, 2 * pipeRoundCount());
      dagColoring = false;
    }
    if (miss & 2) pipeColors = std::min(MAX_COLORS, o[9] + 1);
    bool stale = false;
    if (long(o[12]) > narrowEstimate()) {
      pipeTotEst = o[12];
      stale = true;
    }
    if (pipeEst.size() < size_t(MAX_COLORS)) pipeEst.resize(MAX_COLORS, 0);
    if (pipeEstH.size() < size_t(MAX_COLORS)) pipeEstH.resize(MAX_COLORS, 0);
    for (int c = 0; c < std::min(o[5], PIPE_COLORS); c++) {
      const long cnt = o[17 + c] - o[16 + c], est = pipeEst[c];
      if (cnt > est + est / 4 + 64) {
        pipeEst[c] = cnt;
        stale = true;
      }
      const long hv = o[64 + c], hest = pipeEstH[c];
      if (hv > hest + hest / 4 + 8) {
        pipeEstH[c] = hv;
        stale = true;
      }
    }
    if (stale) dropStepGraphs();
  }

  template <class R, bool EXACT>
  void stepGraphT() {
    // the slot this step will fill holds a step PIPE_SLOTS back: wait for it
    // (at most that many steps in flight) and read it
    while (pipeSerial - pipeTaken >= PIPE_SLOTS) {
      HIPCHECK(hipEventSynchronize(pipeEv[pipeTaken % PIPE_SLOTS]));
      takeSlot(pipeTaken++);
    }
    if (pipeCapacity) {
      pipeCapacity = false;
      throw ApiError(AVBD_E_CAPACITY, "a replayed step exceeded the pair or manifold capacity (the excess was "
                                      "dropped, bits 1 and 2 of the overflow counter); repack to grow it");
    }
    const int E = pipeColorCount(), rounds = pipeRoundCount();
    // the colouring in one launch needs its claim order built outside capture
    const int order = useRank ? AVBD_ORDER_CUSTOM : prm.color_order;
    const bool dag = dagUsable(useUpload, false);
    if (dag) ensureClaimOrder(order);
    const long long key = (1LL << 40) + (long long)E * (1LL << 24) + (long long)rounds * 4 + (dag ? 2 : 0) + cur;
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      const int cur0 = cur, nc0 = numColors;
      numColors = E;
      hipGraph_t g = nullptr;
      HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      graphCapturing = pipeCapturing = true;
      try {
        enqueueStepDevice<R, EXACT>(E, rounds, roundCap <= 0, colorCap <= 0);
      } catch (...) {
        graphCapturing = pipeCapturing = false;
        cur = cur0;
        numColors = nc0;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      graphCapturing = pipeCapturing = false;
      cur = cur0;
      numColors = nc0;
      HIPCHECK(hipStreamEndCapture(stream, &g));
      hipGraphExec_t ex = nullptr;
      const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      if (e != hipSuccess) (void)hipGraphDestroy(g);
      HIPCHECK(e);
      it = graphs.emplace(key, std::make_pair(g, ex)).first;
    }
    HIPCHECK(hipGraphLaunch(it->second.second, stream));
    if (!pipeEv[pipeSerial % PIPE_SLOTS]) HIPCHECK(hipEventCreateWithFlags(&pipeEv[pipeSerial % PIPE_SLOTS],
                                                                            hipEventDisableTiming));
    HIPCHECK(hipEventRecord(pipeEv[pipeSerial % PIPE_SLOTS], stream));
    pipeSerial++;
    cur = 1 - cur;
    steps++;
    pipeUnsettled = true;
  }

  // Bring the host's counts up to the device after graph steps: every call
  // that reads or changes the state does this first.
  void settle() {
    if (!pipeUnsettled) return;
    HIPCHECK(hipStreamSynchronize(stream));
    int64_t last = -1;
    while (pipeTaken < pipeSerial) {
      last = pipeTaken;
      takeSlot(pipeTaken++);
    }
    readScal(0, 10);
    // the lists hold at most pairCap entries (an excess was dropped)
    nPairs = std::min(hscal[0], pairCap);
    nNew = std::min(hscal[1], pairCap);
    nLive = hscal[2];
    contactsTotal = hscal[8];
    if (last >= 0) lastNTot = pipeRing[(last % PIPE_SLOTS) * PIPE_STRIDE + 12];
    HIPCHECK(hipMemcpy(hmeta.data(), cmeta.p, sizeof(int) * hmeta.size(), hipMemcpyDeviceToHost));
    numColors = hmeta[4];
    colorRounds = hmeta[3];
    hcstart.assign(hmeta.begin() + 8, hmeta.begin() + 8 + numColors + 1);
    hheavy.assign(hmeta.begin() + CMETA_HEAVY, hmeta.begin() + CMETA_HEAVY + MAX_COLORS);
    colorFor = -1;  // relax from these colours on a step with host reads
    pipeUnsettled = false;
  }

  // graph steps need a step with host reads since the pack (for the launch
  // shapes) and none of the paths that read on the host inside a step
  bool graphStepOK() const {
    return graphsCfg == AVBD_GRAPHS_STEP && pipeReady && !profiling && !rec && !useUpload && !tuneRequest;
  }

  void step() {
    if (graphStepOK()) {
      switch (prec) {
        case AVBD_PREC_EXACT: stepGraphT<double, true>(); break;
        case AVBD_PREC_F64: stepGraphT<double, false>(); break;
        default: stepGraphT<float, false>(); break;
      }
      return;
    }
    settle();
    if (pipeCapacity) {
      pipeCapacity = false;
      throw ApiError(AVBD_E_CAPACITY, "a replayed step exceeded the pair or manifold capacity (the excess was "
                                      "dropped, bits 1 and 2 of the overflow counter); repack to grow it");
    }
    switch (prec) {
      case AVBD_PREC_EXACT: stepT<double, true>(); break;
      case AVBD_PREC_F64: stepT<double, false>(); break;
      default: stepT<float, false>(); break;
    }
    if (graphsCfg == AVBD_GRAPHS_STEP) {
      // the next step graphs are sized from this step's counts
      pipeShapesFromHost();
      dropStepGraphs();
    }
  }

  // ------------------------------------------------------------- adjoint
  // Recording: the step's contacts (after the narrow phase), its colouring and
  // work lists (after colouring), and a snapshot of the poses and hot records
  // after body initialisation (l = 0) and after each iteration's dual update.
  void recordContacts(int nTot) {
    TapeStep &T = *rec;
    T.nTot = nTot;
    T.nLive = nLive;
    ManStore &m = store[cur];
    T.head.copyFrom(m.head.p, size_t(nTot), stream);
    T.basis.copyFrom((const double *)m.basis.p, size_t(nTot) * BAS, stream);
    T.anchor.copyFrom(m.anchor.p, size_t(nTot) * MAXC * 6, stream);
    T.src.copyFrom(srcSlot.p, size_t(nTot) * MAXC, stream);
    T.live.copyFrom(live[cur].p, size_t(nLive), stream);
  }
  void recordTopology() {
    TapeStep &T = *rec;
    T.numColors = numColors;
    T.color.copyFrom(color[0].p, size_t(N), stream);
    T.cstart.copyFrom(cmeta.p + 8, size_t(numColors + 1), stream);
    T.cwork.copyFrom(cwork.p, size_t(N), stream);
    // as many work items as the step has, counted on the device (no readback)
    T.items.reserve(items.n);
    hipLaunchKernelGGL(k_copy_items, dim3(std::max(1, std::min(blocks(long(items.n)), 1024))), dim3(256), 0, stream,
                       (const int *)(itemStart.p + N), (const uint2 *)items.p, T.items.p, int(items.n));
    T.V.copyFrom(V.p, size_t(N) * VS, stream);
    T.X.reserve(size_t(prm.iterations + 1) * N * XS);
    T.hot.reserve(size_t(prm.iterations + 1) * std::max(T.nTot, 1) * MAXC * HOT);
  }
  // Snapshot l of the poses (X^l: before sweep l, l = 0 after body
  // initialisation) and snapshot h of the hot records (the k, lambda sweep h
  // uses; h = I after the last dual update); -1 skips one.
  void recordSnapshot(int l, int h) {
    TapeStep &T = *rec;
    if (l == 0) {
      T.X0.copyFrom(X0.p, size_t(N) * XS, stream);
      T.Y.copyFrom(Y.p, size_t(N) * XS, stream);
    }
    const size_t nx = size_t(N) * XS, nh = size_t(T.nTot) * MAXC * HOT;
    hipLaunchKernelGGL(k_snapshot, dim3(std::max(1, std::min(blocks(long(std::max(nx, nh))), 1024))), dim3(256), 0,
                       stream, (const double *)X.p, l >= 0 ? nx : size_t(0), l >= 0 ? T.X.p + size_t(l) * nx : nullptr,
                       (const double *)store[cur].hot.p, h >= 0 ? nh : size_t(0),
                       h >= 0 ? T.hot.p + size_t(h) * nh : nullptr);
  }

  // Everything one step hands to the next: poses, velocities, the current
  // contact store and live list, the force lists the broad phase consults,
  // and the colouring.
  void saveCheckpoint(Checkpoint &c) {
    const size_t rs = rsize(), nT = size_t(std::max(lastNTot, 1));
    ManStore &m = store[cur];
    c.X.copyFrom(X.p, size_t(N) * XS, stream);
    c.V.copyFrom(V.p, size_t(N) * VS, stream);
    c.head.copyFrom(m.head.p, nT, stream);
    c.basis.copyFrom(m.basis.p, nT * BAS * rs, stream);
    c.hot.copyFrom(m.hot.p, nT * MAXC * HOT * rs, stream);
    c.Cd.copyFrom(m.Cd.p, nT * MAXC * 4 * rs, stream);
    c.anchor.copyFrom(m.anchor.p, nT * MAXC * 6, stream);
    c.feat.copyFrom(m.feat.p, nT * MAXC, stream);
    c.stick.copyFrom(m.stick.p, nT * MAXC, stream);
    c.live.copyFrom(live[cur].p, size_t(std::max(nLive, 1)), stream);
    c.adjStart.copyFrom(adjStart.p, size_t(N + 1), stream);
    c.adjKey.copyFrom(adjKey.p, adjKey.n, stream);
    c.adjOther.copyFrom(adjOther.p, adjOther.n, stream);
    c.color.copyFrom(color[0].p, size_t(N), stream);
    c.cur = cur;
    c.nLive = nLive;
    c.lastNTot = lastNTot;
    c.steps = steps;
  }
  template <class T> void copyBack(T *dst, const DBuf<T> &src, size_t count) {
    if (count) HIPCHECK(hipMemcpyAsync(dst, src.p, count * sizeof(T), hipMemcpyDeviceToDevice, stream));
  }
  void restoreCheckpoint(const Checkpoint &c) {
    cur = c.cur;
    nLive = c.nLive;
    lastNTot = c.lastNTot;
    steps = c.steps;
    const size_t rs = rsize(), nT = size_t(std::max(lastNTot, 1));
    ManStore &m = store[cur];
    copyBack(X.p, c.X, size_t(N) * XS);
    copyBack(V.p, c.V, size_t(N) * VS);
    copyBack(m.head.p, c.head, nT);
    copyBack(m.basis.p, c.basis, nT * BAS * rs);
    copyBack(m.hot.p, c.hot, nT * MAXC * HOT * rs);
    copyBack(m.Cd.p, c.Cd, nT * MAXC * 4 * rs);
    copyBack(m.anchor.p, c.anchor, nT * MAXC * 6);
    copyBack(m.feat.p, c.feat, nT * MAXC);
    copyBack(m.stick.p, c.stick, nT * MAXC);
    copyBack(live[cur].p, c.live, size_t(std::max(nLive, 1)));
    copyBack(adjStart.p, c.adjStart, size_t(N + 1));
    copyBack(adjKey.p, c.adjKey, adjKey.n);
    copyBack(adjOther.p, c.adjOther, adjOther.n);
    copyBack(color[0].p, c.color, size_t(N));
  }

  void checkAdjointScene() const {
    if (prec == AVBD_PREC_MIXED)
      throw ApiError(AVBD_E_STATE, "the adjoint needs f64 records (precision exact or f64)");
    if (J > 0) throw ApiError(AVBD_E_STATE, "the adjoint covers contacts only; the scene has joints");
    if (prm.post_stabilize) throw ApiError(AVBD_E_STATE, "the adjoint does not cover post-stabilisation");
    if (prm.contact_model) throw ApiError(AVBD_E_STATE, "the adjoint covers the CPU contact law only");
    for (const auto &b : hprops)
      if (b.shape == 1) throw ApiError(AVBD_E_STATE, "the adjoint covers box contacts only; the scene has spheres");
  }

  // Device bytes one recorded step holds.
  static size_t tapeBytes(const TapeStep &T) {
    return T.head.n * sizeof(int4) + T.cwork.n * sizeof(int4) +
           (T.basis.n + T.anchor.n + T.hot.n + T.X.n + T.X0.n + T.Y.n + T.V.n) * sizeof(double) +
           (T.src.n + T.live.n + T.color.n + T.cstart.n) * sizeof(int) + T.items.n * sizeof(uint2);
  }

  // Checkpoints every `every` steps. While the tape recorded so far fits the
  // tape budget, every step is recorded as it runs and the tape stays
  // resident: the backward then reverses it directly instead of re-running
  // each segment from its checkpoint to record it.
  void adjointForward(int nsteps, int every) {
    for (auto &c : ckpts) c.release();
    ckpts.clear();
    ckpts.resize(size_t((nsteps + every - 1) / every));
    tapeResident = adjTapeBudget > 0 && nsteps > 0;
    size_t held = 0;
    struct Reset {
      avbd_ctx *c;
      ~Reset() { c->rec = nullptr; }
    } reset{this};
    if (tapeResident) {
      for (size_t k = size_t(nsteps); k < tape.size(); k++) tape[k].release();
      tape.resize(size_t(nsteps));
    }
    for (int s = 0; s < nsteps; s++) {
      if (s % every == 0) saveCheckpoint(ckpts[size_t(s / every)]);
      rec = tapeResident ? &tape[size_t(s)] : nullptr;
      step();
      rec = nullptr;
      if (tapeResident) {
        // the tape held so far, counted after each step (contacts that appear
        // later make later steps' tapes larger): past the budget it is
        // released and the backward re-runs the segments from their
        // checkpoints, which are kept either way
        held += tapeBytes(tape[size_t(s)]);
        if (double(held) > double(adjTapeBudget)) {
          for (auto &t : tape) t.release();
          tape.clear();
          tapeResident = false;
        }
      }
    }
    saveCheckpoint(endState);
    adjSteps = nsteps;
    adjEvery = every;
  }

  AdjStep adjView(const TapeStep &T) const {
    AdjStep S;
    S.nLive = T.nLive;
    S.numColors = T.numColors;
    S.live = T.live.p; S.head = T.head.p; S.basis = T.basis.p; S.anchor = T.anchor.p; S.src = T.src.p;
    S.hot = T.hot.p; S.X = T.X.p; S.X0 = T.X0.p; S.Y = T.Y.p; S.V = T.V.p;
    S.color = T.color.p; S.cstart = T.cstart.p; S.cwork = T.cwork.p; S.items = T.items.p;
    S.hotStride = size_t(T.nTot) * MAXC * HOT;
    S.xStride = size_t(N) * XS;
    return S;
  }
  AdjBufs adjBufs(int lk) {
    AdjBufs A;
    A.Xb = aXb.p; A.X0b = aX0b.p; A.Yb = aYb.p; A.Vb = aVb.p; A.VbS = aVbS.p;
    A.LK = aLK[lk].p; A.LKprev = aLK[1 - lk].p;
    A.C0b = aC0b.p; A.RWb = aRWb.p; A.MUb = aMUb.p; A.muGrad = aMu.p;
    return A;
  }
  template <class T> void zero(DBuf<T> &b, size_t count) {
    b.reserve(count);
    HIPCHECK(hipMemsetAsync(b.p, 0, sizeof(T) * count, stream));
  }

  // The reverse of one recorded step; lk selects which contact-adjoint buffer
  // holds this step's slots.
  void reverseStep(const TapeStep &T, int &lk, hipStream_t rs) {
    const StepParams P = sp();
    const int I = prm.iterations;
    const AdjStep S = adjView(T);
    const AdjBufs A = adjBufs(lk);
    const bool wide = manifoldsPerBody() > 3;  // many contacts per body: wider teams
    if (nDynamic <= adjSmallMax) {
      // a small scene: the whole reverse step in one workgroup
      if (wide)
        hipLaunchKernelGGL(k_adj_step_small<16>, dim3(1), dim3(ADJ_SMALL_THREADS), 0, rs, N, T.nTot,
                           (const BodyProp *)props.p, S, A, I, P);
      else
        hipLaunchKernelGGL(k_adj_step_small<4>, dim3(1), dim3(ADJ_SMALL_THREADS), 0, rs, N, T.nTot,
                           (const BodyProp *)props.p, S, A, I, P);
      HIPCHECK(hipGetLastError());
      lk = 1 - lk;
      return;
    }
    hipLaunchKernelGGL(k_adj_velocity, dim3(blocks(N)), dim3(256), 0, rs, N, (const BodyProp *)props.p, S, A, I,
                       prm.dt);
    const int gb = std::max(1, std::min(blocks(long(std::max(nDynamic, 1)) * (wide ? 16 : 4)), 2048));
    // each iteration's dual adjoint rides with its first colour in reverse
    for (int l = I - 1; l >= 0; l--) {
      for (int c = T.numColors - 1; c >= 0; c--) {
        if (wide)
          hipLaunchKernelGGL(k_adj_primal<16>, dim3(gb), dim3(256), 0, rs, c, S, A, (const BodyProp *)props.p, l,
                             P);
        else
          hipLaunchKernelGGL(k_adj_primal<4>, dim3(gb), dim3(256), 0, rs, c, S, A, (const BodyProp *)props.p, l,
                             P);
      }
    }
    if (T.nLive > 0)
      hipLaunchKernelGGL(k_adj_narrow, dim3(blocks(T.nLive, 128)), dim3(128), 0, rs, S, A,
                         (const BodyProp *)props.p, P);
    hipLaunchKernelGGL(k_adj_init, dim3(blocks(N)), dim3(256), 0, rs, N, (const BodyProp *)props.p, S, A, P);
    if (T.nTot) {
      const size_t slots = size_t(T.nTot) * MAXC;
      HIPCHECK(hipMemsetAsync(aLK[lk].p, 0, sizeof(double) * slots * 6, rs));
      HIPCHECK(hipMemsetAsync(aC0b.p, 0, sizeof(double) * slots * 3, rs));
      HIPCHECK(hipMemsetAsync(aRWb.p, 0, sizeof(double) * slots * 6, rs));
      HIPCHECK(hipMemsetAsync(aMUb.p, 0, sizeof(double) * T.nTot, rs));
    }
    HIPCHECK(hipGetLastError());
    lk = 1 - lk;
  }

  double manifoldsPerBody() const { return double(std::max(lastNTot, 0)) / double(std::max(nDynamic, 1)); }

  void adjointBackward(const double *dLdPos, double *dLdMu, double *dLdV0) {
    if (adjSteps <= 0) throw ApiError(AVBD_E_STATE, "no adjoint forward recorded");
    const size_t slots = size_t(manCap) * MAXC;
    zero(aXb, size_t(N) * XS); zero(aX0b, size_t(N) * XS); zero(aYb, size_t(N) * XS);
    zero(aVb, size_t(N) * VS); zero(aVbS, size_t(N) * VS);
    zero(aLK[0], slots * 6); zero(aLK[1], slots * 6); zero(aC0b, slots * 3); zero(aRWb, slots * 6);
    zero(aMUb, size_t(manCap)); zero(aMu, size_t(N));
    {
      std::vector<double> h(size_t(N) * XS, 0.0);
      for (int i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) h[size_t(i) * XS + k] = dLdPos[i * 3 + k];
      HIPCHECK(hipMemcpyAsync(aXb.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, stream));
      HIPCHECK(hipStreamSynchronize(stream));
    }
    struct Reset {
      avbd_ctx *c;
      ~Reset() { c->rec = nullptr; }
    } reset{this};
    int lk = 0;
    if (tapeResident) {
      for (int k = adjSteps - 1; k >= 0; k--) reverseStep(tape[size_t(k)], lk, stream);
    }
    // Segment re-runs, pipelined: segment seg is re-run on stream into half
    // seg % 2 of the segment tape, then reversed on rstream while segment
    // seg - 1 is re-run into the other half (which the reverse of segment
    // seg + 1 has finished reading). The reverse reads only the tape, the
    // bodies' properties and the adjoint buffers, none of which a re-run
    // writes; the segments are still reversed in order, last to first.
    const int nseg = tapeResident ? 0 : int(ckpts.size());
    const size_t E = size_t(adjEvery);
    if (nseg > 0 && tape.size() < 2 * E) tape.resize(2 * E);
    auto rerun = [&](int seg) {
      const size_t half = size_t(seg % 2) * E;
      restoreCheckpoint(ckpts[size_t(seg)]);
      const int len = std::min(adjEvery, adjSteps - seg * adjEvery);
      for (int k = 0; k < len; k++) {
        rec = &tape[half + size_t(k)];
        step();
      }
      rec = nullptr;
      HIPCHECK(hipEventRecord(evRerun, stream));
    };
    if (nseg > 0) rerun(nseg - 1);
    for (int seg = nseg - 1; seg >= 0; seg--) {
      const size_t half = size_t(seg % 2) * E;
      const int len = std::min(adjEvery, adjSteps - seg * adjEvery);
      HIPCHECK(hipStreamWaitEvent(rstream, evRerun, 0));
      for (int k = len - 1; k >= 0; k--) reverseStep(tape[half + size_t(k)], lk, rstream);
      HIPCHECK(hipEventRecord(evRev[seg % 2], rstream));
      if (seg > 0) {
        // the half segment seg - 1 records into was last read by the reverse
        // of segment seg + 1 (its buffers may be regrown on the host)
        if (seg + 1 < nseg) HIPCHECK(hipEventSynchronize(evRev[(seg + 1) % 2]));
        rerun(seg - 1);
      }
    }
    if (nseg > 0) HIPCHECK(hipStreamWaitEvent(stream, evRev[0], 0));
    restoreCheckpoint(endState);
    std::vector<double> mu(static_cast<size_t>(N)), vb(static_cast<size_t>(N) * VS);
    HIPCHECK(hipMemcpyAsync(mu.data(), aMu.p, sizeof(double) * N, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipMemcpyAsync(vb.data(), aVb.p, sizeof(double) * N * VS, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    for (int i = 0; i < N; i++) {
      if (dLdMu) dLdMu[i] = mu[size_t(i)];
      // pack sets the previous velocity equal to the velocity
      if (dLdV0)
        for (int k = 0; k < 3; k++) dLdV0[i * 3 + k] = vb[size_t(i) * VS + k] + vb[size_t(i) * VS + 6 + k];
    }
  }
};

namespace {

int fail(avbd_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

// Every call but avbd_step first brings the host's counts up to date after
// replayed steps (settleFirst).
template <class F> int guarded(avbd_ctx *c, F &&f, bool settleFirst = true) {
  if (!c) return AVBD_E_INVALID;
  try {
    if (settleFirst) c->settle();
    f();
    return AVBD_OK;
  } catch (const ApiError &e) {
    return fail(c, e.code, e.what());
  } catch (const HipError &e) {
    return fail(c, AVBD_E_HIP, e.what());
  } catch (const std::exception &e) {
    return fail(c, AVBD_E_STATE, e.what());
  }
}

// The scene's contact material (avbd_scene static_friction .. rest_offset),
// validated as Rigid.setMaterial validates it (rigid.js:144-214), into
// c.hmatRaw: 5 numbers per body, offsets NaN where the solver's apply.
void readMaterials(avbd_ctx &c, const avbd_scene &s) {
  const int N = s.num_bodies;
  c.hmatRaw.assign(size_t(N) * 5, 0.0);
  const bool given = s.static_friction || s.dynamic_friction || s.restitution || s.contact_offset || s.rest_offset;
  auto nonNeg = [](double v) { return std::isfinite(v) && v >= 0; };
  for (int i = 0; i < N; i++) {
    double *r = &c.hmatRaw[size_t(i) * 5];
    r[0] = s.static_friction ? s.static_friction[i] : s.friction[i];
    r[1] = s.dynamic_friction ? s.dynamic_friction[i] : s.friction[i];
    r[2] = s.restitution ? s.restitution[i] : 0.0;
    r[3] = s.contact_offset ? s.contact_offset[i] : std::nan("");
    r[4] = s.rest_offset ? s.rest_offset[i] : std::nan("");
    if (!given) continue;
    const std::string at = "body " + std::to_string(i) + ": ";
    if (!nonNeg(r[0]) || !nonNeg(r[1]))
      throw ApiError(AVBD_E_INVALID, at + "static and dynamic friction must be finite and non-negative");
    if (r[1] > r[0]) throw ApiError(AVBD_E_INVALID, at + "dynamic friction must not exceed static friction");
    if (!(r[2] >= 0 && r[2] <= 1)) throw ApiError(AVBD_E_INVALID, at + "restitution must be in [0, 1]");
    if (!std::isnan(r[3]) && !nonNeg(r[3]))
      throw ApiError(AVBD_E_INVALID, at + "contact offset must be finite and non-negative");
    if (std::isinf(r[4])) throw ApiError(AVBD_E_INVALID, at + "rest offset must be finite");
    if (!std::isnan(r[3]) && !std::isnan(r[4]) && r[4] > r[3])
      throw ApiError(AVBD_E_INVALID, at + "rest offset must not exceed contact offset");
  }
}

// The material resolved against the solver's offsets as
// bodyContactSignature does (backend.js:156-185): contact offset at least 0,
// rest offset at most the contact offset. Uploads it with the pair test's
// contact offsets and sets the broad-phase grid cell to at least twice the
// largest reach of a grid body: its radius plus, under the GPU contact model,
// the largest contact offset of any body (a pair within reach of each other
// then has its centres in neighbouring cells).
void resolveMaterials(avbd_ctx &c) {
  const int N = c.N;
  std::vector<BodyMat> M(static_cast<size_t>(N));
  std::vector<double> off(static_cast<size_t>(N));
  double offMax = 0;
  for (int i = 0; i < N; i++) {
    const double *r = &c.hmatRaw[size_t(i) * 5];
    BodyMat &m = M[size_t(i)];
    m.staticFriction = r[0];
    m.dynamicFriction = r[1];
    m.restitution = r[2];
    m.contactOffset = jmax(0.0, std::isnan(r[3]) ? c.prm.contact_offset : r[3]);
    m.restOffset = jmin(std::isnan(r[4]) ? c.prm.rest_offset : r[4], m.contactOffset);
    m.pad = 0;
    off[size_t(i)] = m.contactOffset;
    offMax = std::max(offMax, m.contactOffset);
  }
  if (c.stream) HIPCHECK(hipStreamSynchronize(c.stream));  // no step may be reading the old values
  HIPCHECK(hipMemcpy(c.mat.p, M.data(), sizeof(BodyMat) * N, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c.skinOff.p, off.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  const double reach = c.prm.contact_model ? c.rmaxSmall + offMax : c.rmaxSmall;
  c.invh = 1.0 / std::max(2.0 * reach * (1.0 + 1e-9), 1e-6);
}

void packScene(avbd_ctx &c, const avbd_scene &s) {
  if (s.num_bodies <= 0) throw ApiError(AVBD_E_INVALID, "num_bodies must be positive");
  if (!s.shape || !s.size || !s.density || !s.friction || !s.pos)
    throw ApiError(AVBD_E_INVALID, "shape, size, density, friction and pos are required");
  const int N = s.num_bodies;
  c.releaseAll();
  HIPCHECK(hipSetDevice(c.device));
  HIPCHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  HIPCHECK(hipStreamCreateWithFlags(&c.rstream, hipStreamNonBlocking));
  HIPCHECK(hipEventCreateWithFlags(&c.evRerun, hipEventDisableTiming));
  HIPCHECK(hipEventCreateWithFlags(&c.evRev[0], hipEventDisableTiming));
  HIPCHECK(hipEventCreateWithFlags(&c.evRev[1], hipEventDisableTiming));
  HIPCHECK(hipHostMalloc((void **)&c.hscal, sizeof(int) * 16, hipHostMallocDefault));
  HIPCHECK(hipHostMalloc((void **)&c.pipeRing, sizeof(int) * c.PIPE_SLOTS * PIPE_STRIDE, hipHostMallocDefault));
  std::memset(c.pipeRing, 0, sizeof(int) * c.PIPE_SLOTS * PIPE_STRIDE);
  HIPCHECK(hipHostGetDevicePointer((void **)&c.pipeRingDev, c.pipeRing, 0));
  c.pipeSerial = c.pipeTaken = 0;
  c.pipeReady = c.pipeUnsettled = c.pipeCapacity = false;
  c.pipeColors = c.pipeRounds = 0;
  c.pipeEst.clear();
  c.pipeEstH.clear();
  c.pipeTotEst = 0;
  if (s.num_joints < 0) throw ApiError(AVBD_E_INVALID, "num_joints must not be negative");
  if (s.num_joints > 0 && (!s.joint_a || !s.joint_b || !s.joint_ra || !s.joint_rb || !s.joint_stiff_lin ||
                           !s.joint_stiff_ang || !s.joint_fracture))
    throw ApiError(AVBD_E_INVALID, "joint arrays are required when num_joints > 0");
  c.N = N;
  c.J = s.num_joints;
  c.steps = 0;
  c.nLive = 0;
  c.cur = 0;
  c.nPairs = c.nNew = 0;
  c.numColors = 0;
  c.colorFor = -1;

  // scenes must be contiguous runs of bodies
  std::vector<int> sceneOf(N, 0);
  if (s.scene_of) {
    for (int i = 0; i < N; i++) {
      sceneOf[i] = s.scene_of[i];
      if (sceneOf[i] < 0 || (i > 0 && sceneOf[i] < sceneOf[i - 1]))
        throw ApiError(AVBD_E_INVALID, "scene_of must be non-negative and non-decreasing");
    }
  }
  const int S = sceneOf[N - 1] + 1;
  c.S = S;

  // mass properties (rigid.js:29-133)
  std::vector<BodyProp> P(N);
  std::vector<double> X(size_t(N) * XS, 0.0), V(size_t(N) * VS, 0.0);
  for (int i = 0; i < N; i++) {
    BodyProp &b = P[i];
    const double sx = s.size[i * 3], sy = s.size[i * 3 + 1], sz = s.size[i * 3 + 2];
    b.sx = sx; b.sy = sy; b.sz = sz;
    b.hx = sx * 0.5; b.hy = sy * 0.5; b.hz = sz * 0.5;
    b.friction = s.friction[i];
    b.shape = s.shape[i];
    b.scene = sceneOf[i];
    b.pad = 0;
    const double rho = s.density[i];
    if (b.shape == AVBD_SHAPE_SPHERE) {
      const double r = sx * 0.5;
      b.radius = r;
      b.mass = (4.0 / 3.0) * M_PI * r * r * r * rho;
      const double I = (2.0 / 5.0) * b.mass * r * r;
      b.m0 = b.m1 = b.m2 = I;
    } else if (b.shape == AVBD_SHAPE_BOX) {
      b.mass = sx * sy * sz * rho;
      b.m0 = ((sy * sy + sz * sz) / 12.0) * b.mass;
      b.m1 = ((sx * sx + sz * sz) / 12.0) * b.mass;
      b.m2 = ((sx * sx + sy * sy) / 12.0) * b.mass;
      const double hx = sx * 0.5, hy = sy * 0.5, hz = sz * 0.5;
      b.radius = std::sqrt(hx * hx + hy * hy + hz * hz);
    } else {
      throw ApiError(AVBD_E_INVALID, "unknown shape");
    }
    double *x = &X[size_t(i) * XS];
    for (int k = 0; k < 3; k++) x[k] = s.pos[i * 3 + k];
    if (s.quat) for (int k = 0; k < 4; k++) x[3 + k] = s.quat[i * 4 + k];
    else { x[3] = x[4] = x[5] = 0; x[6] = 1; }
    double *v = &V[size_t(i) * VS];
    for (int k = 0; k < 3; k++) {
      v[k] = s.vel ? s.vel[i * 3 + k] : 0.0;
      v[3 + k] = s.angvel ? s.angvel[i * 3 + k] : 0.0;
      v[6 + k] = v[k];  // prevVelocityLin starts equal to the velocity
    }
  }
  // large bodies per scene (radius > 8x the scene's mean radius), grid cell
  std::vector<double> rsum(S, 0.0);
  std::vector<int> scnt(S, 0);
  for (int i = 0; i < N; i++) { rsum[P[i].scene] += P[i].radius; scnt[P[i].scene]++; }
  std::vector<int> sceneStart(S + 1, 0), sls(S + 1, 0), slist, largeList;
  for (int i = 0; i < N; i++) sceneStart[P[i].scene + 1]++;
  for (int k = 0; k < S; k++) sceneStart[k + 1] += sceneStart[k];
  double rmax = 0;
  for (int i = 0; i < N; i++) {
    const int sc = P[i].scene;
    const double thr = std::max(rsum[sc] / scnt[sc] * 8.0, 1e-6);
    P[i].large = P[i].radius > thr ? 1 : 0;
    if (!P[i].large) rmax = std::max(rmax, P[i].radius);
  }
  for (int sc = 0; sc < S; sc++) {
    sls[sc] = int(slist.size());
    for (int i = sceneStart[sc]; i < sceneStart[sc + 1]; i++)
      if (P[i].large) { slist.push_back(i); largeList.push_back(i); }
  }
  sls[S] = int(slist.size());
  c.nLarge = int(largeList.size());
  c.rmaxSmall = rmax;
  readMaterials(c, s);
  uint32_t tsz = 64;
  while (tsz < uint32_t(2 * N)) tsz <<= 1;
  c.tableMask = tsz - 1;

  c.hprops = P;
  c.mat.reserve(N);
  c.skinOff.reserve(N);
  resolveMaterials(c);  // and the grid cell
  c.nDynamic = 0;
  for (const BodyProp &b : P) c.nDynamic += b.mass > 0;
  c.props.reserve(N);
  HIPCHECK(hipMemcpy(c.props.p, P.data(), sizeof(BodyProp) * N, hipMemcpyHostToDevice));
  c.X.reserve(size_t(N) * XS); c.X0.reserve(size_t(N) * XS); c.Y.reserve(size_t(N) * XS);
  c.V.reserve(size_t(N) * VS);
  HIPCHECK(hipMemcpy(c.X.p, X.data(), sizeof(double) * X.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c.X0.p, X.data(), sizeof(double) * X.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c.Y.p, X.data(), sizeof(double) * X.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c.V.p, V.data(), sizeof(double) * V.size(), hipMemcpyHostToDevice));
  c.DQ.reserve(size_t(N) * 8 * c.rsize());
  HIPCHECK(hipMemset(c.DQ.p, 0, size_t(N) * 8 * c.rsize()));

  c.cellOf.reserve(N); c.entries.reserve(N); c.entryCell.reserve(N);
  c.bpScratch.reserve(size_t(N) * BP_CAP);
  c.bucketCnt.reserve(tsz); c.bucketStart.reserve(tsz + 1); c.bucketCursor.reserve(tsz);
  c.sceneLargeStart.reserve(S + 1); c.sceneLarges.reserve(slist.size() + 1);
  c.sceneStart.reserve(S + 1); c.largeList.reserve(largeList.size() + 1);
  HIPCHECK(hipMemcpy(c.sceneLargeStart.p, sls.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice));
  if (!slist.empty()) HIPCHECK(hipMemcpy(c.sceneLarges.p, slist.data(), sizeof(int) * slist.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c.sceneStart.p, sceneStart.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice));
  c.hsceneStart = sceneStart;
  if (!largeList.empty())
    HIPCHECK(hipMemcpy(c.largeList.p, largeList.data(), sizeof(int) * largeList.size(), hipMemcpyHostToDevice));

  // an overflow seen since the last pack regrows the arena (backend.js:
  // 1683-1739: report, then regrow on the next repack)
  c.pairCap = std::max(c.pairCapCfg > 0 ? c.pairCapCfg : std::max(32 * N, 4096), c.needPairs + c.needPairs / 4);
  c.manCap = std::max(c.manCapCfg > 0 ? c.manCapCfg : std::max(32 * N, 4096), c.needMan + c.needMan / 4);
  c.needPairs = c.needMan = 0;
  c.countRev.reserve(N); c.startRev.reserve(N + 1);
  c.countNewRev.reserve(N); c.startNewRev.reserve(N + 1);
  c.pairs.reserve(c.pairCap);
  c.newPairs.reserve(c.pairCap);
  c.store[0].reserve(c.manCap, c.rsize());
  c.store[1].reserve(c.manCap, c.rsize());
  c.live[0].reserve(c.manCap); c.live[1].reserve(c.manCap);
  c.keep.reserve(c.manCap); c.keepPos.reserve(c.manCap + 1);
  c.srcSlot.reserve(size_t(c.manCap) * MAXC);
  c.lastNTot = 0;
  c.adjSteps = 0;
  c.stateWritten = false;

  // joints (joint.js:78-113)
  std::vector<JointDev> JD(c.J);
  for (int j = 0; j < c.J; j++) {
    JointDev &d = JD[j];
    std::memset(&d, 0, sizeof(d));
    d.a = s.joint_a[j];
    d.b = s.joint_b[j];
    if (d.b < 0 || d.b >= N || d.a < -1 || d.a >= N || d.a == d.b)
      throw ApiError(AVBD_E_INVALID, "joint " + std::to_string(j) + " has invalid bodies");
    if (d.a >= 0 && sceneOf[d.a] != sceneOf[d.b]) throw ApiError(AVBD_E_INVALID, "joint crosses scenes");
    for (int k = 0; k < 3; k++) { d.rA[k] = s.joint_ra[j * 3 + k]; d.rB[k] = s.joint_rb[j * 3 + k]; }
    d.stiffLin = s.joint_stiff_lin[j];
    d.stiffAng = s.joint_stiff_ang[j];
    d.fracture = s.joint_fracture[j];
    d.kind = s.joint_kind ? s.joint_kind[j] : AVBD_JOINT_BALL;
    if (d.kind != AVBD_JOINT_BALL && d.kind != AVBD_JOINT_SPRING)
      throw ApiError(AVBD_E_INVALID, "joint " + std::to_string(j) + " has an unknown kind");
    if (d.kind == AVBD_JOINT_SPRING) {
      if (d.a < 0) throw ApiError(AVBD_E_INVALID, "spring " + std::to_string(j) + " needs two bodies");
      if (d.stiffAng < 0) {
        // spring.js:78-84: rest length = the anchors' current distance
        double pa[3], pb[3], dd[3];
        xform(pa, &X[size_t(d.a) * XS], &X[size_t(d.a) * XS + 3], d.rA);
        xform(pb, &X[size_t(d.b) * XS], &X[size_t(d.b) * XS + 3], d.rB);
        for (int k = 0; k < 3; k++) dd[k] = pa[k] - pb[k];
        d.stiffAng = std::sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
      }
    }
    const double ax = d.a >= 0 ? s.size[d.a * 3] : 0, ay = d.a >= 0 ? s.size[d.a * 3 + 1] : 0,
                 az = d.a >= 0 ? s.size[d.a * 3 + 2] : 0;
    const double sx = ax + s.size[d.b * 3], sy = ay + s.size[d.b * 3 + 1], sz = az + s.size[d.b * 3 + 2];
    d.torqueArm = sx * sx + sy * sy + sz * sz;
  }
  c.hjointKind.resize(size_t(c.J));
  for (int j = 0; j < c.J; j++) c.hjointKind[size_t(j)] = JD[size_t(j)].kind;
  // the one-launch solve of a batch: joints grouped by scene in scene order
  {
    std::vector<int> sjs(S + 1, 0);
    bool grouped = true;
    int last = 0;
    for (int j = 0; j < c.J; j++) {
      const int sc = sceneOf[JD[size_t(j)].b];
      if (sc < last) grouped = false;
      last = sc;
      sjs[sc + 1]++;
    }
    for (int k = 0; k < S; k++) sjs[k + 1] += sjs[k];
    int biggest = 0;
    for (int k = 0; k < S; k++) biggest = std::max(biggest, sceneStart[k + 1] - sceneStart[k]);
    c.sceneSolveOK = S >= 2 && grouped && biggest <= 4096;
    c.sceneJointStart.reserve(S + 1);
    HIPCHECK(hipMemcpy(c.sceneJointStart.p, sjs.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice));
  }
  c.joints.reserve(std::max(c.J, 1));
  if (c.J) HIPCHECK(hipMemcpy(c.joints.p, JD.data(), sizeof(JointDev) * c.J, hipMemcpyHostToDevice));

  c.deg.reserve(N); c.adjStart.reserve(N + 1); c.adjCursor.reserve(N);
  c.adjKey.reserve(2 * (size_t(c.manCap) + c.J) + 1);
  c.adjOther.reserve(2 * (size_t(c.manCap) + c.J) + 1);
  c.adjKeyFill.reserve(2 * (size_t(c.manCap) + c.J) + 1);
  c.adjOtherFill.reserve(2 * (size_t(c.manCap) + c.J) + 1);
  c.adjOwner.reserve(2 * (size_t(c.manCap) + c.J) + 1);
  c.color[0].reserve(N); c.color[1].reserve(N);
  std::vector<int> col(N);
  for (int i = 0; i < N; i++) col[i] = P[i].mass > 0 ? 0 : -1;
  HIPCHECK(hipMemcpy(c.color[0].p, col.data(), sizeof(int) * N, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c.color[1].p, col.data(), sizeof(int) * N, hipMemcpyHostToDevice));
  c.colorCur = 0;
  c.cstart.reserve(MAX_COLORS + 1);
  c.cbodies.reserve(N);
  c.cwork.reserve(N);
  c.hwork.reserve(N);
  c.itemCnt.reserve(N); c.itemStart.reserve(N + 1); c.itemJoints.reserve(N);
  c.chist.reserve(size_t(blocks(N)) * MAX_COLORS); c.coffs.reserve(size_t(blocks(N)) * MAX_COLORS);
  c.ctotal.reserve(MAX_COLORS);
  c.claimRank.reserve(N); c.uploadColor.reserve(N);
  c.useRank = c.useUpload = false;
  c.items.reserve(2 * (size_t(MAXC) * c.manCap + c.J) + 1);
  c.scal.reserve(16);
  c.pc.reserve(PC_N);
  // block sums of the largest scan a step makes (the manifold keep flags of a
  // step replayed whole are scanned at capacity)
  c.scanTmp[0].reserve(size_t(std::max({c.manCap, int(c.tableMask) + 1, N + 1}) + 1023) / 1024 + 1);
  c.cbytes.reserve(2 * MAX_COLORS);
  c.cmeta.reserve(CMETA_N);
  HIPCHECK(hipMemset(c.scal.p, 0, sizeof(int) * 16));
  HIPCHECK(hipMemset(c.pc.p, 0, sizeof(int) * PC_N));
  c.evs.resize(AVBD_NUM_PHASES);
  for (auto &e : c.evs) HIPCHECK(hipEventCreate(&e));
  // force lists of the initial scene (joints only) for the first broad phase
  c.buildAdjacency();
  HIPCHECK(hipStreamSynchronize(c.stream));
  c.packed = true;
}

}  // namespace

// State carried over a repack: the previous scene's device buffers, taken out
// of the context before it packs the new scene, freed when this goes.
struct RepackCarry {
  DBuf<double> X, X0, V;
  ManStore store;
  DBuf<int> live, color;
  int nLive = 0;
  ~RepackCarry() {
    X.release(); X0.release(); V.release(); store.release(); live.release(); color.release();
  }
};

template <class R>
int carryManifolds(avbd_ctx &c, RepackCarry &old, const int *remap) {
  if (old.nLive == 0) return 0;
  c.keep.reserve(size_t(old.nLive));
  c.keepPos.reserve(size_t(old.nLive) + 1);
  hipLaunchKernelGGL(k_repack_keep, dim3(blocks(old.nLive)), dim3(256), 0, c.stream, old.nLive, old.live.p,
                     (const int4 *)old.store.head.p, remap, c.keep.p);
  c.scan(c.keep.p, c.keepPos.p, old.nLive, c.scal.p + 2);
  c.readScal(2, 1);
  const int kept = c.hscal[2];
  if (kept > c.manCap)
    throw ApiError(AVBD_E_CAPACITY, "manifolds " + std::to_string(kept) + " exceed capacity " + std::to_string(c.manCap));
  hipLaunchKernelGGL((k_repack_manifolds<R>), dim3(old.nLive), dim3(64), 0, c.stream, (const int *)old.live.p,
                     (const int *)c.keep.p, (const int *)c.keepPos.p, remap, old.store.view<R>(),
                     c.store[c.cur].view<R>(), c.live[c.cur].p);
  HIPCHECK(hipGetLastError());
  return kept;
}

extern "C" {

int avbd_abi_version(void) { return AVBD_ABI_VERSION; }

void avbd_default_params(avbd_params *p) {
  if (!p) return;
  p->dt = 1.0 / 60.0;
  p->gravity = -10.0;
  p->beta_lin = 10000.0;
  p->beta_ang = 100.0;
  p->alpha = 0.95;
  p->gamma = 0.999;
  p->penalty_min = 1.0;
  p->penalty_max = 1.0e10;
  p->collision_margin = 0.01;
  p->stick_threshold = 0.00001;
  p->iterations = 10;
  p->post_stabilize = 0;
  p->rotated_inertia = 1;
  p->cached_contact_jacobians = 1;
  p->color_order = AVBD_ORDER_INDEX;
  p->contact_model = AVBD_CONTACTS_CPU;
  p->contact_offset = 0.005;
  p->rest_offset = -0.005;
  p->restitution_threshold = 1.0;
  p->contact_persistence = 0.002;
}

int avbd_create(avbd_ctx **out, const avbd_config *cfg) {
  if (!out) return AVBD_E_INVALID;
  *out = nullptr;
  avbd_ctx *c = new (std::nothrow) avbd_ctx();
  if (!c) return AVBD_E_STATE;
  avbd_default_params(&c->prm);
  if (cfg) {
    c->device = cfg->device;
    c->prec = cfg->precision;
    c->team = cfg->team > 0 ? cfg->team : 16;
    c->exactTeam = cfg->team != 1;
    c->manCapCfg = cfg->manifold_cap;
    c->graphsCfg = cfg->graphs;
    c->adjSmallMax = cfg->adjoint_small == 0 ? 64 : cfg->adjoint_small;
    c->adjTapeBudget = cfg->adjoint_tape_mb == 0 ? (4096L << 20)
                       : cfg->adjoint_tape_mb < 0 ? -1L
                                                  : long(cfg->adjoint_tape_mb) << 20;
    c->pairCapCfg = cfg->pair_cap;
    c->sceneSolveCfg = cfg->scene_solve;
  }
  if (c->prec < AVBD_PREC_EXACT || c->prec > AVBD_PREC_MIXED) {
    delete c;
    return AVBD_E_INVALID;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->device) {
    delete c;
    return AVBD_E_HIP;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && cus > 0)
    c->simds = 4 * cus;
  // the static staging array's address on this device, for the entry point
  // that allocates nothing
  c->poseScratch = staticPoseScratch(c->device);
  if (!c->poseScratch) {
    delete c;
    return AVBD_E_HIP;
  }
  *out = c;
  return AVBD_OK;
}

int avbd_destroy(avbd_ctx *ctx) {
  if (!ctx) return AVBD_E_INVALID;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return AVBD_OK;
}

const char *avbd_last_error(const avbd_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int avbd_set_params(avbd_ctx *ctx, const avbd_params *p) {
  return guarded(ctx, [&] {
    if (!p) throw ApiError(AVBD_E_INVALID, "null params");
    if (!(p->dt > 0) || p->iterations < 1) throw ApiError(AVBD_E_INVALID, "dt must be > 0 and iterations >= 1");
    if (p->color_order != AVBD_ORDER_INDEX && p->color_order != AVBD_ORDER_PRIORITY)
      throw ApiError(AVBD_E_INVALID, "color_order must be AVBD_ORDER_INDEX or AVBD_ORDER_PRIORITY");
    if (!p->cached_contact_jacobians)
      throw ApiError(AVBD_E_INVALID, "only cached contact Jacobians are implemented on the GPU");
    if (p->contact_model != AVBD_CONTACTS_CPU && p->contact_model != AVBD_CONTACTS_GPU)
      throw ApiError(AVBD_E_INVALID, "contact_model must be AVBD_CONTACTS_CPU or AVBD_CONTACTS_GPU");
    if (!std::isfinite(p->contact_offset) || !std::isfinite(p->rest_offset) ||
        !std::isfinite(p->restitution_threshold) || !(p->contact_persistence >= 0) ||
        !std::isfinite(p->contact_persistence))
      throw ApiError(AVBD_E_INVALID, "contact offsets, restitution threshold and persistence must be finite "
                                     "(persistence >= 0)");
    ctx->prm = *p;
    ctx->dropAdjoint();
    ctx->dropGraphs();  // parameters are baked into the captured launches
    if (ctx->packed) resolveMaterials(*ctx);  // offsets and the grid follow the solver's
  });
}

int avbd_pack(avbd_ctx *ctx, const avbd_scene *scene) {
  return guarded(ctx, [&] {
    if (!scene) throw ApiError(AVBD_E_INVALID, "null scene");
    ctx->packed = false;
    packScene(*ctx, *scene);
  });
}

int avbd_repack(avbd_ctx *ctx, const avbd_scene *scene, const int32_t *old_index) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed (use avbd_pack)");
    if (!scene || !old_index) throw ApiError(AVBD_E_INVALID, "null scene or old_index");
    avbd_ctx &c = *ctx;
    const int oldN = c.N, N = scene->num_bodies;
    if (N <= 0) throw ApiError(AVBD_E_INVALID, "num_bodies must be positive");
    std::vector<int> remap(oldN, -1);
    int last = -1;
    for (int i = 0; i < N; i++) {
      const int o = old_index[i];
      if (o < -1 || o >= oldN) throw ApiError(AVBD_E_INVALID, "old_index out of range");
      if (o < 0) continue;
      if (o <= last) throw ApiError(AVBD_E_INVALID, "surviving bodies must keep their order");
      last = o;
      remap[o] = i;
    }
    // the joints: exactly the survivors of the previous scene's, in order
    HIPCHECK(hipStreamSynchronize(c.stream));
    std::vector<JointDev> oldJ(c.J);
    if (c.J) HIPCHECK(hipMemcpy(oldJ.data(), c.joints.p, sizeof(JointDev) * c.J, hipMemcpyDeviceToHost));
    std::vector<JointDev> keptJ;
    for (const JointDev &j : oldJ) {
      const int na = j.a < 0 ? -1 : remap[j.a], nb = remap[j.b];
      if (nb < 0 || (j.a >= 0 && na < 0)) continue;
      JointDev k = j;
      k.a = na;
      k.b = nb;
      keptJ.push_back(k);
    }
    if (scene->num_joints > 0 && (!scene->joint_a || !scene->joint_b))
      throw ApiError(AVBD_E_INVALID, "joint arrays are required when num_joints > 0");
    if (scene->num_joints != int(keptJ.size()))
      throw ApiError(AVBD_E_INVALID, "a repack keeps the surviving joints and adds none (" +
                                         std::to_string(keptJ.size()) + " survive, scene has " +
                                         std::to_string(scene->num_joints) + "); add joints with avbd_pack");
    for (int j = 0; j < scene->num_joints; j++)
      if (scene->joint_a[j] != keptJ[j].a || scene->joint_b[j] != keptJ[j].b ||
          (scene->joint_kind ? scene->joint_kind[j] : AVBD_JOINT_BALL) != keptJ[j].kind)
        throw ApiError(AVBD_E_INVALID, "joint " + std::to_string(j) + " is not the surviving joint in that place");

    RepackCarry old;
    std::swap(old.X, c.X);
    std::swap(old.X0, c.X0);
    std::swap(old.V, c.V);
    std::swap(old.live, c.live[c.cur]);
    std::swap(old.color, c.color[0]);
    std::swap(old.store.head, c.store[c.cur].head);
    std::swap(old.store.basis, c.store[c.cur].basis);
    std::swap(old.store.hot, c.store[c.cur].hot);
    std::swap(old.store.Cd, c.store[c.cur].Cd);
    std::swap(old.store.anchor, c.store[c.cur].anchor);
    std::swap(old.store.feat, c.store[c.cur].feat);
    std::swap(old.store.stick, c.store[c.cur].stick);
    old.nLive = c.nLive;

    c.packed = false;
    packScene(c, *scene);  // fresh buffers, the description's state everywhere
    DBuf<int> dOld, dRemap;
    dOld.reserve(N);
    dRemap.reserve(oldN);
    HIPCHECK(hipMemcpy(dOld.p, old_index, sizeof(int) * N, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dRemap.p, remap.data(), sizeof(int) * oldN, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_repack_bodies, dim3(blocks(N)), dim3(256), 0, c.stream, N, (const int *)dOld.p,
                       (const BodyProp *)c.props.p, (const double *)old.X.p, (const double *)old.X0.p,
                       (const double *)old.V.p, (const int *)old.color.p, c.X.p, c.X0.p, c.V.p, c.color[0].p);
    HIPCHECK(hipGetLastError());
    c.nLive = c.prec == AVBD_PREC_MIXED ? carryManifolds<float>(c, old, dRemap.p)
                                        : carryManifolds<double>(c, old, dRemap.p);
    if (c.J) HIPCHECK(hipMemcpy(c.joints.p, keptJ.data(), sizeof(JointDev) * c.J, hipMemcpyHostToDevice));
    // the force lists the next broad phase tells existing pairs by
    c.buildAdjacency();
    HIPCHECK(hipStreamSynchronize(c.stream));
    // survivors carry device state, not the description's
    c.stateWritten = true;
    dOld.release();
    dRemap.release();
  });
}

int avbd_step(avbd_ctx *ctx, int32_t nsteps) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    HIPCHECK(hipSetDevice(ctx->device));
    if (nsteps > 0) ctx->dropAdjoint();
    for (int s = 0; s < nsteps; s++) ctx->step();
  }, false);
}

int avbd_sync(avbd_ctx *ctx) {
  return guarded(ctx, [&] {
    if (ctx->stream) HIPCHECK(hipStreamSynchronize(ctx->stream));
  });
}

int avbd_read_state(avbd_ctx *ctx, double *pos, double *quat, double *vel, double *angvel, double *prev_vel,
                    int32_t n) {
  return guarded(ctx, [&] {
    if (!ctx->packed || n != ctx->N) throw ApiError(AVBD_E_INVALID, "body count mismatch");
    std::vector<double> X(size_t(n) * XS), V(size_t(n) * VS);
    HIPCHECK(hipMemcpyAsync(X.data(), ctx->X.p, sizeof(double) * X.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(V.data(), ctx->V.p, sizeof(double) * V.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; i++) {
      const double *x = &X[size_t(i) * XS], *v = &V[size_t(i) * VS];
      for (int k = 0; k < 3; k++) {
        if (pos) pos[i * 3 + k] = x[k];
        if (vel) vel[i * 3 + k] = v[k];
        if (angvel) angvel[i * 3 + k] = v[3 + k];
        if (prev_vel) prev_vel[i * 3 + k] = v[6 + k];
      }
      if (quat) for (int k = 0; k < 4; k++) quat[i * 4 + k] = x[3 + k];
    }
  });
}

int avbd_export_poses(avbd_ctx *ctx, double alpha, float *dst, int32_t n, int32_t dst_on_device) {
  return guarded(ctx, [&] {
    if (!ctx->packed || n != ctx->N) throw ApiError(AVBD_E_INVALID, "body count mismatch");
    if (!dst) throw ApiError(AVBD_E_INVALID, "dst is NULL");
    // Math.min(1, Math.max(0, alpha)), and a NaN (which fails a < 1) draws
    // the current pose, as the renderer does (renderer.js:1329-1330)
    const double a = std::isnan(alpha) ? 1.0 : std::min(1.0, std::max(0.0, alpha));
    if (n == 0) return;
    float4 *out;
    if (dst_on_device) {
      hipPointerAttribute_t attr;
      if (hipPointerGetAttributes(&attr, dst) != hipSuccess) {
        (void)hipGetLastError();
        throw ApiError(AVBD_E_INVALID, "dst is not device memory");
      }
      if (attr.type != hipMemoryTypeDevice || attr.device != ctx->device)
        throw ApiError(AVBD_E_INVALID, "dst is not device memory on the context's device");
      if (reinterpret_cast<uintptr_t>(dst) % 16 != 0) throw ApiError(AVBD_E_INVALID, "dst must be 16-byte aligned");
      out = reinterpret_cast<float4 *>(dst);
    } else {
      ctx->poseStage.reserve(size_t(n) * 2);
      out = ctx->poseStage.p;
    }
    // algorithmic bytes: p and q of one or two poses read (56 B each), 32 B written
    const int lerp = a < 1.0 ? 1 : 0;
    ctx->timedLaunch(AVBD_K_POSES, double(n) * (lerp ? 144 : 88), k_export_poses, dim3(blocks(n)), dim3(256), n, a,
                     lerp, reinterpret_cast<const double2 *>(ctx->X.p),
                     reinterpret_cast<const double2 *>(ctx->X0.p), out);
    HIPCHECK(hipGetLastError());
    if (!dst_on_device) {
      HIPCHECK(hipMemcpyAsync(dst, out, sizeof(float4) * 2 * size_t(n), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
  });
}

int avbd_export_poses_static(avbd_ctx *ctx, double alpha, float *dst, int32_t n) {
  return guarded(ctx, [&] {
    if (!ctx->packed || n != ctx->N) throw ApiError(AVBD_E_INVALID, "body count mismatch");
    if (!dst) throw ApiError(AVBD_E_INVALID, "dst is NULL");
    if (n > AVBD_STATIC_POSE_BODIES)
      throw ApiError(AVBD_E_INVALID, "avbd_export_poses_static takes at most " +
                                         std::to_string(AVBD_STATIC_POSE_BODIES) + " bodies (" + std::to_string(n) +
                                         " packed); use avbd_export_poses");
    if (!ctx->poseScratch) throw ApiError(AVBD_E_STATE, "no static staging address for this device");
    const double a = std::isnan(alpha) ? 1.0 : std::min(1.0, std::max(0.0, alpha));
    if (n == 0) return;
    // the kernel writes all 2n float4 of the staging prefix the copy reads,
    // so nothing needs clearing first; the launch and the copy are ordered on
    // the context's stream
    const int lerp = a < 1.0 ? 1 : 0;
    hipLaunchKernelGGL(k_export_poses, dim3(blocks(n)), dim3(256), 0, ctx->stream, n, a, lerp,
                       reinterpret_cast<const double2 *>(ctx->X.p), reinterpret_cast<const double2 *>(ctx->X0.p),
                       ctx->poseScratch);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(dst, ctx->poseScratch, sizeof(float4) * 2 * size_t(n), hipMemcpyDeviceToHost,
                            ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  });
}

int avbd_pose_source(avbd_ctx *ctx, const double **pose, const double **start_pose, int32_t *stride_bytes) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    if (pose) *pose = ctx->X.p;
    if (start_pose) *start_pose = ctx->X0.p;
    if (stride_bytes) *stride_bytes = int32_t(sizeof(double) * XS);
  });
}

int avbd_write_state(avbd_ctx *ctx, const double *pos, const double *quat, const double *vel, const double *angvel,
                     const double *prev_vel, int32_t n) {
  return guarded(ctx, [&] {
    if (!ctx->packed || n != ctx->N) throw ApiError(AVBD_E_INVALID, "body count mismatch");
    ctx->dropAdjoint();
    ctx->stateWritten = true;
    std::vector<double> X(size_t(n) * XS), V(size_t(n) * VS);
    HIPCHECK(hipMemcpyAsync(X.data(), ctx->X.p, sizeof(double) * X.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(V.data(), ctx->V.p, sizeof(double) * V.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; i++) {
      double *x = &X[size_t(i) * XS], *v = &V[size_t(i) * VS];
      for (int k = 0; k < 3; k++) {
        if (pos) x[k] = pos[i * 3 + k];
        if (vel) v[k] = vel[i * 3 + k];
        if (angvel) v[3 + k] = angvel[i * 3 + k];
        if (prev_vel) v[6 + k] = prev_vel[i * 3 + k];
      }
      if (quat) for (int k = 0; k < 4; k++) x[3 + k] = quat[i * 4 + k];
    }
    HIPCHECK(hipMemcpyAsync(ctx->X.p, X.data(), sizeof(double) * X.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipMemcpyAsync(ctx->V.p, V.data(), sizeof(double) * V.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  });
}

int avbd_read_pairs(avbd_ctx *ctx, int32_t *out, int32_t cap, int32_t *count) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    if (count) *count = ctx->nPairs;
    const int n = std::min(cap, ctx->nPairs);
    if (out && n > 0) {
      HIPCHECK(hipMemcpyAsync(out, ctx->pairs.p, sizeof(int2) * n, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
  });
}

int avbd_read_coloring(avbd_ctx *ctx, int32_t *color_of, int32_t n, int32_t *num_colors) {
  return guarded(ctx, [&] {
    if (!ctx->packed || n != ctx->N) throw ApiError(AVBD_E_INVALID, "body count mismatch");
    if (num_colors) *num_colors = ctx->numColors;
    if (color_of) {
      HIPCHECK(hipMemcpyAsync(color_of, ctx->color[ctx->colorCur].p, sizeof(int) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
      HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
  });
}

static std::vector<JointDev> hostJoints(avbd_ctx *ctx) {
  std::vector<JointDev> J(ctx->J);
  if (ctx->J) {
    HIPCHECK(hipMemcpyAsync(J.data(), ctx->joints.p, sizeof(JointDev) * ctx->J, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  }
  return J;
}

int avbd_read_joints(avbd_ctx *ctx, double *out, int32_t n) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    if (n != ctx->J) throw ApiError(AVBD_E_INVALID, "joint count mismatch");
    if (n == 0) return;
    if (!out) throw ApiError(AVBD_E_INVALID, "out is NULL");
    const std::vector<JointDev> J = hostJoints(ctx);
    for (int j = 0; j < n; j++) {
      const JointDev &d = J[j];
      double *o = out + size_t(j) * 16;
      for (int k = 0; k < 16; k++) o[k] = 0.0;
      o[12] = d.broken ? 1.0 : 0.0;
      if (d.kind == AVBD_JOINT_SPRING) {
        o[13] = d.kLin[0];
        o[14] = d.CLin[0];
        o[15] = d.lamLin[0];
      } else {
        for (int k = 0; k < 3; k++) {
          o[k] = d.kLin[k]; o[3 + k] = d.kAng[k]; o[6 + k] = d.lamLin[k]; o[9 + k] = d.lamAng[k];
        }
      }
    }
  });
}

int avbd_set_joint_anchor(avbd_ctx *ctx, int32_t joint, const double *rA) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    if (joint < 0 || joint >= ctx->J) throw ApiError(AVBD_E_INVALID, "joint index out of range");
    if (ctx->hjointKind[size_t(joint)] != AVBD_JOINT_BALL) throw ApiError(AVBD_E_INVALID, "not a ball joint");
    if (!rA) throw ApiError(AVBD_E_INVALID, "rA is NULL");
    ctx->dropAdjoint();
    // stream-ordered: the next step reads the new anchor
    ctx->anchorStage.resize(size_t(ctx->J) * 3);
    double *h = &ctx->anchorStage[size_t(joint) * 3];
    for (int k = 0; k < 3; k++) h[k] = rA[k];
    HIPCHECK(hipMemcpyAsync(reinterpret_cast<char *>(ctx->joints.p + joint) + offsetof(JointDev, rA), h,
                            sizeof(double) * 3, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  });
}

int avbd_read_step_start(avbd_ctx *ctx, double *pos, double *quat, int32_t n) {
  return guarded(ctx, [&] {
    if (!ctx->packed || n != ctx->N) throw ApiError(AVBD_E_INVALID, "body count mismatch");
    std::vector<double> X(size_t(n) * XS);
    HIPCHECK(hipMemcpyAsync(X.data(), ctx->X0.p, sizeof(double) * X.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; i++) {
      if (pos) for (int k = 0; k < 3; k++) pos[i * 3 + k] = X[size_t(i) * XS + k];
      if (quat) for (int k = 0; k < 4; k++) quat[i * 4 + k] = X[size_t(i) * XS + 3 + k];
    }
  });
}

int avbd_read_forces(avbd_ctx *ctx, int32_t *kabn, uint32_t *features, int32_t cap, int32_t *count) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    std::vector<JointDev> J = hostJoints(ctx);
    std::vector<int> live(ctx->nLive);
    std::vector<int4> head(ctx->nLive);
    if (ctx->nLive) {
      HIPCHECK(hipMemcpyAsync(live.data(), ctx->live[ctx->cur].p, sizeof(int) * ctx->nLive, hipMemcpyDeviceToHost,
                              ctx->stream));
      HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
    std::vector<int4> allHead(ctx->nLive ? size_t(*std::max_element(live.begin(), live.end())) + 1 : 0);
    std::vector<uint32_t> allFeat(allHead.size() * MAXC);
    if (!allHead.empty()) {
      HIPCHECK(hipMemcpyAsync(allHead.data(), ctx->store[ctx->cur].head.p, sizeof(int4) * allHead.size(),
                              hipMemcpyDeviceToHost, ctx->stream));
      HIPCHECK(hipMemcpyAsync(allFeat.data(), ctx->store[ctx->cur].feat.p, sizeof(uint32_t) * allFeat.size(),
                              hipMemcpyDeviceToHost, ctx->stream));
      HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
    int k = 0;
    auto put = [&](int kind, int a, int b, int nc, const uint32_t *f) {
      if (k < cap) {
        if (kabn) { kabn[k * 4] = kind; kabn[k * 4 + 1] = a; kabn[k * 4 + 2] = b; kabn[k * 4 + 3] = nc; }
        if (features)
          for (int s = 0; s < MAXC; s++) features[k * MAXC + s] = (f && s < nc) ? f[s] : 0u;
      }
      k++;
    };
    for (int j = 0; j < ctx->J; j++)
      if (J[j].broken != 2) put(J[j].kind == AVBD_JOINT_SPRING ? 2 : 1, J[j].a, J[j].b, 0, nullptr);
    for (int t = 0; t < ctx->nLive; t++) {
      const int4 h = allHead[live[t]];
      put(0, h.x, h.y, h.z, &allFeat[size_t(live[t]) * MAXC]);
    }
    if (count) *count = k;
  });
}

int avbd_read_manifold(avbd_ctx *ctx, int32_t index, double *basis_mu, double *contacts, int32_t *num_contacts) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    std::vector<JointDev> J = hostJoints(ctx);
    int nj = 0;
    for (auto &j : J) nj += j.broken != 2;
    const int t = index - nj;
    if (t < 0 || t >= ctx->nLive) throw ApiError(AVBD_E_INVALID, "index is not a live manifold");
    int m;
    HIPCHECK(hipMemcpyAsync(&m, ctx->live[ctx->cur].p + t, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ManStore &st = ctx->store[ctx->cur];
    const size_t rs = ctx->rsize();
    int4 h;
    std::vector<char> bas(BAS * rs), hot(MAXC * HOT * rs);
    std::vector<double> anc(MAXC * 6);
    std::vector<int32_t> stick(MAXC);
    HIPCHECK(hipMemcpyAsync(&h, st.head.p + m, sizeof(int4), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(bas.data(), st.basis.p + size_t(m) * BAS * rs, BAS * rs, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(hot.data(), st.hot.p + size_t(m) * MAXC * HOT * rs, MAXC * HOT * rs, hipMemcpyDeviceToHost,
                            ctx->stream));
    HIPCHECK(hipMemcpyAsync(anc.data(), st.anchor.p + size_t(m) * MAXC * 6, sizeof(double) * MAXC * 6,
                            hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(stick.data(), st.stick.p + size_t(m) * MAXC, sizeof(int32_t) * MAXC,
                            hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    auto rd = [&](const std::vector<char> &v, size_t i) {
      return rs == sizeof(float) ? double(((const float *)v.data())[i]) : ((const double *)v.data())[i];
    };
    if (basis_mu) for (int k = 0; k < 10; k++) basis_mu[k] = rd(bas, k);
    if (num_contacts) *num_contacts = h.z;
    if (contacts)
      for (int s = 0; s < h.z; s++) {
        double *o = contacts + s * 16;
        for (int k = 0; k < 6; k++) o[k] = anc[s * 6 + k];
        for (int k = 0; k < 3; k++) {
          o[6 + k] = rd(hot, size_t(s) * HOT + HC0 + k);
          o[9 + k] = rd(hot, size_t(s) * HOT + HK + k);
          o[12 + k] = rd(hot, size_t(s) * HOT + HL + k);
        }
        o[15] = stick[s];
      }
  });
}

int avbd_read_manifolds(avbd_ctx *ctx, int32_t *head, double *basis_mu, double *contacts, int32_t cap,
                        int32_t contact_cap, int32_t *count, int32_t *num_contacts) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    const int nl = ctx->nLive;
    if (count) *count = nl;
    std::vector<int> live(size_t(std::max(nl, 0)));
    if (nl) {
      HIPCHECK(hipMemcpyAsync(live.data(), ctx->live[ctx->cur].p, sizeof(int) * nl, hipMemcpyDeviceToHost,
                              ctx->stream));
      HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
    const size_t span = nl ? size_t(*std::max_element(live.begin(), live.end())) + 1 : 0;
    ManStore &st = ctx->store[ctx->cur];
    const size_t rs = ctx->rsize();
    std::vector<int4> h(span);
    if (span) {
      HIPCHECK(hipMemcpyAsync(h.data(), st.head.p, sizeof(int4) * span, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
    long nc = 0;
    for (int t = 0; t < nl; t++) nc += h[size_t(live[t])].z;
    if (num_contacts) *num_contacts = int32_t(nc);
    if (!head && !basis_mu && !contacts) return;
    std::vector<char> bas, hot;
    std::vector<double> anc;
    std::vector<int32_t> stick;
    if (span && basis_mu) {
      bas.resize(span * BAS * rs);
      HIPCHECK(hipMemcpyAsync(bas.data(), st.basis.p, bas.size(), hipMemcpyDeviceToHost, ctx->stream));
    }
    if (span && contacts) {
      hot.resize(span * MAXC * HOT * rs);
      anc.resize(span * MAXC * 6);
      stick.resize(span * MAXC);
      HIPCHECK(hipMemcpyAsync(hot.data(), st.hot.p, hot.size(), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHECK(hipMemcpyAsync(anc.data(), st.anchor.p, sizeof(double) * anc.size(), hipMemcpyDeviceToHost,
                              ctx->stream));
      HIPCHECK(hipMemcpyAsync(stick.data(), st.stick.p, sizeof(int32_t) * stick.size(), hipMemcpyDeviceToHost,
                              ctx->stream));
    }
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    auto rd = [&](const std::vector<char> &v, size_t i) {
      return rs == sizeof(float) ? double(((const float *)v.data())[i]) : ((const double *)v.data())[i];
    };
    long k = 0;
    for (int t = 0; t < nl && t < cap; t++) {
      const size_t m = size_t(live[t]);
      const int4 hd = h[m];
      if (head) { head[t * 4] = hd.x; head[t * 4 + 1] = hd.y; head[t * 4 + 2] = hd.z; head[t * 4 + 3] = hd.w; }
      if (basis_mu)
        for (int e = 0; e < 10; e++) basis_mu[size_t(t) * 10 + e] = rd(bas, m * BAS + e);
      if (contacts)
        for (int sl = 0; sl < hd.z && k < contact_cap; sl++, k++) {
          double *o = contacts + size_t(k) * 16;
          const size_t si = m * MAXC + size_t(sl);
          for (int e = 0; e < 6; e++) o[e] = anc[si * 6 + e];
          for (int e = 0; e < 3; e++) {
            o[6 + e] = rd(hot, si * HOT + HC0 + e);
            o[9 + e] = rd(hot, si * HOT + HK + e);
            o[12 + e] = rd(hot, si * HOT + HL + e);
          }
          o[15] = stick[si];
        }
    }
  });
}

int avbd_read_contact_model(avbd_ctx *ctx, double *material, double *bias, int32_t cap, int32_t contact_cap) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    if (cap < 0 || contact_cap < 0) throw ApiError(AVBD_E_INVALID, "negative capacity");
    const int nl = ctx->nLive;
    if (nl == 0 || (!material && !bias)) return;
    std::vector<int> live(static_cast<size_t>(nl));
    HIPCHECK(hipMemcpyAsync(live.data(), ctx->live[ctx->cur].p, sizeof(int) * nl, hipMemcpyDeviceToHost,
                            ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    const size_t span = size_t(*std::max_element(live.begin(), live.end())) + 1;
    ManStore &st = ctx->store[ctx->cur];
    const size_t rs = ctx->rsize();
    std::vector<int4> h(span);
    std::vector<char> bas(span * BAS * rs), hot(span * MAXC * HOT * rs);
    HIPCHECK(hipMemcpyAsync(h.data(), st.head.p, sizeof(int4) * span, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(bas.data(), st.basis.p, bas.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(hot.data(), st.hot.p, hot.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    auto rd = [&](const std::vector<char> &v, size_t i) {
      return rs == sizeof(float) ? double(((const float *)v.data())[i]) : ((const double *)v.data())[i];
    };
    long k = 0;
    for (int t = 0; t < nl && t < cap; t++) {
      const size_t m = size_t(live[t]);
      if (material) {
        material[size_t(t) * 3] = rd(bas, m * BAS + BMU);
        material[size_t(t) * 3 + 1] = rd(bas, m * BAS + BMUD);
        material[size_t(t) * 3 + 2] = rd(bas, m * BAS + BCD);
      }
      for (int sl = 0; sl < h[m].z && k < contact_cap; sl++, k++)
        if (bias) bias[k] = rd(hot, (m * MAXC + size_t(sl)) * HOT + HB);
    }
  });
}

int avbd_read_counters(avbd_ctx *ctx, avbd_counters *out) {
  return guarded(ctx, [&] {
    if (!out) throw ApiError(AVBD_E_INVALID, "null output");
    std::memset(out, 0, sizeof(*out));
    out->steps = ctx->steps;
    out->bodies = ctx->N;
    int nd = 0;
    for (auto &p : ctx->hprops) nd += p.mass > 0;
    out->dynamic_bodies = nd;
    out->pairs = ctx->nPairs;
    out->new_pairs = ctx->nNew;
    out->manifolds = ctx->nLive;
    out->colors = ctx->numColors;
    out->color_rounds = ctx->colorRounds;
    if (ctx->packed) {
      HIPCHECK(hipStreamSynchronize(ctx->stream));
      std::vector<JointDev> J = hostJoints(ctx);
      int nj = 0;
      for (auto &j : J) nj += j.broken != 2;
      out->joints = nj;
      int ov = 0;
      HIPCHECK(hipMemcpy(&ov, ctx->scal.p + 5, sizeof(int), hipMemcpyDeviceToHost));
      out->overflow = ov;
      // contacts = sum of nc over live manifolds
      if (ctx->nLive) {
        std::vector<int> live(ctx->nLive);
        HIPCHECK(hipMemcpy(live.data(), ctx->live[ctx->cur].p, sizeof(int) * ctx->nLive, hipMemcpyDeviceToHost));
        const int mx = *std::max_element(live.begin(), live.end()) + 1;
        std::vector<int4> head(mx);
        HIPCHECK(hipMemcpy(head.data(), ctx->store[ctx->cur].head.p, sizeof(int4) * mx, hipMemcpyDeviceToHost));
        int nc = 0;
        for (int t : live) nc += head[t].z;
        out->contacts = nc;
      }
    }
  });
}

int avbd_set_profiling(avbd_ctx *ctx, int32_t on) {
  return guarded(ctx, [&] {
    if (ctx->stream) HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->harvest();
    ctx->profiling = on != 0;
    for (double &m : ctx->profMs) m = 0;
    ctx->profSteps = 0;
    for (int k = 0; k < AVBD_NUM_KERNELS; k++) ctx->kMs[k] = ctx->kBytes[k] = 0, ctx->kLaunches[k] = 0;
  });
}

int avbd_autotune(avbd_ctx *ctx, int32_t *team, float *ms) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    if (ctx->prec == AVBD_PREC_EXACT) throw ApiError(AVBD_E_STATE, "the exact precision has a fixed team of 1");
    HIPCHECK(hipSetDevice(ctx->device));
    ctx->dropAdjoint();
    ctx->tuneRequest = true;
    ctx->step();
    if (team) *team = ctx->useScenes() ? ctx->sceneTeamShape() : ctx->team;
    if (ms) for (int k = 0; k < 6; k++) ms[k] = ctx->tuneMs[k];
  });
}

int avbd_read_launch_shape(avbd_ctx *ctx, int32_t *team, int32_t *block, float *grid) {
  return guarded(ctx, [&] {
    const bool scenes = ctx->packed && ctx->useScenes();
    const bool xt = ctx->prec == AVBD_PREC_EXACT;
    if (team) *team = scenes ? ctx->sceneTeamShape() : xt ? (ctx->exactTeam ? XT_TEAM : 1) : ctx->team;
    if (block) *block = scenes ? ctx->sceneBlockShape() : xt ? ctx->exactBlock : ctx->sweepBlock;
    if (grid) std::memcpy(grid, ctx->tuneGrid, sizeof(ctx->tuneGrid));
  });
}

int avbd_set_launch_shape(avbd_ctx *ctx, int32_t team, int32_t block) {
  return guarded(ctx, [&] {
    if (ctx->prec == AVBD_PREC_EXACT) {
      // the exact team (XT_TEAM lanes per body) is fixed; its workgroup is not
      if (team != XT_TEAM || (block != 64 && block != 128 && block != 256))
        throw ApiError(AVBD_E_INVALID, "the exact precision takes team 32 and block 64, 128 or 256");
      ctx->exactBlock = block;
    } else if (ctx->packed && ctx->useScenes()) {
      if ((team != 4 && team != 8 && team != 16) || (block != 128 && block != 256 && block != 512))
        throw ApiError(AVBD_E_INVALID, "a batch solved in one launch takes team 4, 8 or 16 and block 128, 256 or 512");
      ctx->sceneTeam = team;
      ctx->sceneBlock = block;
    } else {
      if ((team != 1 && team != 4 && team != 8 && team != 16 && team != 32 && team != 64) ||
          (block != 64 && block != 128 && block != 256))
        throw ApiError(AVBD_E_INVALID, "team must be 1, 4, 8, 16, 32 or 64 and block 64, 128 or 256");
      ctx->team = team;
      ctx->sweepBlock = block;
      ctx->pipeReady = false;  // the heavy-body counts under the new team: read on the next step
    }
    ctx->dropAdjoint();
    ctx->dropGraphs();  // captured with the old shapes
  });
}

int avbd_set_body_order(avbd_ctx *ctx, const int32_t *rank, int32_t n) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    ctx->dropAdjoint();
    ctx->dropGraphs();  // a replayed step's colouring is captured with its claim order
    if (!rank) {
      ctx->useRank = false;
      return;
    }
    if (n != ctx->N) throw ApiError(AVBD_E_INVALID, "rank must hold one entry per body");
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(hipMemcpyAsync(ctx->claimRank.p, rank, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->hrank.assign(rank, rank + n);
    ctx->corderFor = -1;  // the one-launch colouring's order follows the ranks
    ctx->useRank = true;
  });
}

int avbd_set_coloring_limits(avbd_ctx *ctx, int32_t rounds, int32_t colors) {
  return guarded(ctx, [&] {
    if (rounds < 0 || colors < 0 || colors > MAX_COLORS)
      throw ApiError(AVBD_E_INVALID, "rounds must be >= 0 and colors in [0, 256]");
    ctx->roundCap = rounds;
    ctx->colorCap = colors;
    ctx->dropAdjoint();
    ctx->dropGraphs();  // a replayed step's colouring is captured with its limits
  });
}

int avbd_read_inertial_target(avbd_ctx *ctx, double *pos, double *quat, int32_t n) {
  return guarded(ctx, [&] {
    if (!ctx->packed || n != ctx->N) throw ApiError(AVBD_E_INVALID, "body count mismatch");
    std::vector<double> Y(size_t(n) * XS);
    HIPCHECK(hipMemcpyAsync(Y.data(), ctx->Y.p, sizeof(double) * Y.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; i++) {
      if (pos) for (int k = 0; k < 3; k++) pos[i * 3 + k] = Y[size_t(i) * XS + k];
      if (quat) for (int k = 0; k < 4; k++) quat[i * 4 + k] = Y[size_t(i) * XS + 3 + k];
    }
  });
}

int avbd_upload_coloring(avbd_ctx *ctx, const int32_t *color_of, int32_t n) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    ctx->dropAdjoint();
    if (!color_of) {
      ctx->useUpload = false;
      return;
    }
    if (n != ctx->N) throw ApiError(AVBD_E_INVALID, "color_of must hold one entry per body");
    for (int i = 0; i < n; i++)
      if (ctx->hprops[size_t(i)].mass > 0 && (color_of[i] < 0 || color_of[i] >= MAX_COLORS))
        throw ApiError(AVBD_E_INVALID, "a dynamic body's colour is outside [0, 256)");
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(hipMemcpyAsync(ctx->uploadColor.p, color_of, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->useUpload = true;
  });
}

int avbd_adjoint_forward(avbd_ctx *ctx, int32_t nsteps, int32_t ckpt_every) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    if (nsteps <= 0 || ckpt_every <= 0) throw ApiError(AVBD_E_INVALID, "nsteps and ckpt_every must be positive");
    ctx->checkAdjointScene();
    // the gradient with respect to the initial velocity is that of the
    // velocity pack set (and the previous velocity pack set equal to it)
    if (ctx->steps != 0 || ctx->stateWritten)
      throw ApiError(AVBD_E_STATE, "the adjoint forward starts from the packed scene: pack it again first");
    HIPCHECK(hipSetDevice(ctx->device));
    ctx->adjointForward(nsteps, ckpt_every);
  });
}

int avbd_adjoint_backward(avbd_ctx *ctx, const double *dL_dpos, double *dL_dmu, double *dL_dv0) {
  return guarded(ctx, [&] {
    if (!ctx->packed) throw ApiError(AVBD_E_STATE, "no scene packed");
    if (!dL_dpos) throw ApiError(AVBD_E_INVALID, "dL_dpos is required");
    HIPCHECK(hipSetDevice(ctx->device));
    ctx->adjointBackward(dL_dpos, dL_dmu, dL_dv0);
  });
}

int avbd_read_kernel_stats(avbd_ctx *ctx, int32_t kernel, double *ms_total, int64_t *launches, double *bytes_total) {
  return guarded(ctx, [&] {
    if (kernel < 0 || kernel >= AVBD_NUM_KERNELS) throw ApiError(AVBD_E_INVALID, "unknown kernel id");
    if (ctx->stream) HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->harvest();
    if (ms_total) *ms_total = ctx->kMs[kernel];
    if (launches) *launches = ctx->kLaunches[kernel];
    if (bytes_total) *bytes_total = ctx->kBytes[kernel];
  });
}

int avbd_read_profile(avbd_ctx *ctx, float *ms, int32_t n, int32_t *steps) {
  return guarded(ctx, [&] {
    if (steps) *steps = ctx->profSteps;
    for (int k = 0; k < n && k < AVBD_NUM_PHASES; k++)
      ms[k] = ctx->profSteps ? float(ctx->profMs[k] / ctx->profSteps) : 0.0f;
  });
}

}  // extern "C"
