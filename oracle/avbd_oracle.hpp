// avbd_oracle.hpp — CPU f64 restatement of the reference AVBD rigid-body step.
//
// TEST INFRASTRUCTURE. This is the oracle: the HIP library is checked against
// it, and bench.py times it as the CPU baseline. Nothing in the product path
// links or calls it.
//
// It restates the reference's f64 CPU engine (src/math/maths.js,
// src/physics/{solver,rigid,force,manifold,collide,broadphase,coloring,joint}.js)
// with every floating-point operation in the reference's order, so that, built
// with -ffp-contract=off, it reproduces the reference's trajectories bit for
// bit. Per-function citations are in avbd_oracle.cpp.
#pragma once
#include <cstdint>
#include <limits>
#include <vector>

namespace orc {

constexpr int MAX_CONTACTS = 8;  // collide.js:22

struct Params {                  // solver.js:92-236 (defaultParams)
  double dt = 1.0 / 60.0;
  double gravity = -10.0;
  int iterations = 10;
  double betaLin = 10000.0;
  double betaAng = 100.0;
  double alpha = 0.95;
  double gamma = 0.999;
  double penaltyMin = 1.0;
  double penaltyMax = 1.0e10;
  double collisionMargin = 0.01;
  double stickThreshold = 0.00001;
  bool postStabilize = false;
  bool rotatedInertia = true;
  bool cachedContactJacobians = true;
  // The reference GPU backend's contact model (shaders.js:440-454,
  // 821-1055, 1105-1367, 1777-1850, 2229-2264; solver.js:142-162;
  // rigid.js:144-214): 0 = the CPU engine's (the default), 1 = speculative
  // contact skins, rest offsets, restitution, static and dynamic friction.
  int contactModel = 0;
  double contactOffset = 0.005;        // solver.js:151
  double restOffset = -0.005;          // solver.js:152
  double restitutionThreshold = 1.0;   // solver.js:155
  double contactPersistence = 0.002;   // solver.js:162
};

enum Shape : int { BOX = 0, SPHERE = 1 };

struct Body {                    // rigid.js:29-98
  int shape = BOX;
  double size[3];
  double density, friction, mass, moment[3], radius;
  // contact material of the GPU contact model (Rigid.setMaterial,
  // rigid.js:144-214); offsets NaN = the solver's
  double staticFriction = 0, dynamicFriction = 0, restitution = 0;
  double contactOffset = std::numeric_limits<double>::quiet_NaN();
  double restOffset = std::numeric_limits<double>::quiet_NaN();
  double p[3], q[4];             // positionLin / positionAng
  double p0[3], q0[4];           // initialLin / initialAng (x_t)
  double py[3], qy[4];           // inertialLin / inertialAng (y)
  double v[3], w[3], vprev[3];   // velocityLin / velocityAng / prevVelocityLin
  std::vector<int> forces;       // force ids touching the body, creation order
};

struct Contact {                 // manifold.js:34-70
  uint32_t feature = 0;
  double rA[3] = {0, 0, 0}, rB[3] = {0, 0, 0};
  double C0[3] = {0, 0, 0}, k[3] = {0, 0, 0}, lam[3] = {0, 0, 0}, C[3] = {0, 0, 0};
  bool stick = false;
  double bias = 0.0;             // GPU contact model: one-step restitution target
  void reset() {
    feature = 0;
    for (int i = 0; i < 3; i++) k[i] = lam[i] = C0[i] = C[i] = 0.0;
    stick = false;
    bias = 0.0;
  }
};

enum ForceKind : int { MANIFOLD = 0, JOINT = 1, SPRING = 2 };

struct Force {
  int kind = MANIFOLD;
  int a = -1, b = -1;            // bodyA / bodyB (a == -1: world-anchored joint)
  int64_t seq = 0;               // creation order (position in the force list)
  // --- manifold (manifold.js:165-199)
  int nc = 0;
  double basis[9] = {0};
  double mu = 0.0;               // CPU model: sqrt(mu_A mu_B); GPU model: the static coefficient
  double muD = 0.0;              // GPU model: the dynamic coefficient
  double contactDistance = 0.0;  // GPU model: the pair's contact skin
  Contact c[MAX_CONTACTS];
  double jac[MAX_CONTACTS * 18] = {0};
  bool jacValid = false;
  // --- joint (joint.js:78-129)
  double jrA[3] = {0}, jrB[3] = {0};
  double stiffLin = 0, stiffAng = 0, fracture = 0, torqueArm = 0;
  bool broken = false;
  double C0Lin[3] = {0}, C0Ang[3] = {0}, kLin[3] = {0}, kAng[3] = {0};
  double lamLin[3] = {0}, lamAng[3] = {0}, CLin[3] = {0}, CAng[3] = {0};
  // --- spring (spring.js:63-105); anchors in jrA / jrB, torn flag in broken
  double stiffness = 0, rest = 0, tearStrain = 0, penalty = 0, springC = 0, peakStrain = 0;
};

// Primal sweep schedule.
enum Schedule : int {
  SCHED_REVERSE = 0,   // reference default: bodies n-1 ... 0 (solver.js:419-422)
  SCHED_ORDER = 1,     // caller-supplied bodyOrder (solver.js:225-233)
  SCHED_COLOR = 2,     // greedy colouring of the force graph taken right after
                       // force initialisation; bodies in colour order
};

struct Coloring {                // coloring.js:43-118
  std::vector<int> colorOf;      // -1 for static bodies
  std::vector<int> offsets;      // numColors + 1
  std::vector<int> entries;      // dynamic bodies grouped by colour, index order
  int numColors = 0;
};

struct World {
  Params prm;
  std::vector<Body> bodies;
  std::vector<Force> pool;       // every force's record (ids index here)
  std::vector<int> freeForces;   // pool slots of dropped contact manifolds, reused by newForce
  std::vector<int> forces;       // live force ids, creation order
  int64_t nextSeq = 0;
  Schedule schedule = SCHED_REVERSE;
  std::vector<int> bodyOrder;    // for SCHED_ORDER
  Coloring coloring;             // last colouring (SCHED_COLOR)
  std::vector<int> lastPairs;    // broad-phase pairs of the last step (i, j)...
  bool recordPairs = false;
  int broadphase = 0;            // 0 grid (exactly the brute-force pair set), 1 brute force
  int colorOrder = 0;            // 0 body index, 1 lowbias32 priority
  // GPU contact model: contacts warm-started from an old contact with another
  // feature key (the geometric persistence match), counted over all steps
  int64_t geometricMatches = 0;

  int addBox(const double size[3], double density, double friction,
             const double pos[3], const double vel[3]);
  int addSphere(double diameter, double density, double friction,
                const double pos[3], const double vel[3]);
  int addJoint(int a, int b, const double rA[3], const double rB[3],
               double stiffLin, double stiffAng, double fracture);
  // A distance spring (spring.js:63-105); rest < 0 takes the current length.
  int addSpring(int a, int b, const double rA[3], const double rB[3], double stiffness, double rest,
                double tearStrain);
  int removeBodies(const std::vector<char> &doomed);
  void step();
  void buildColoring(Coloring &out) const;
};

// lowbias32, the colour-claim priority of the reference GPU (layout.js:50-57).
uint32_t colorPriority(uint32_t x);

// Narrow phase on its own, for tests: fills up to MAX_CONTACTS contacts
// (feature, rA, rB) and the basis; returns the contact count.
// skin: the GPU contact model's contact distance; mids, when given, receives
// each contact's world midpoint (xA + xB) / 2 (3 per contact).
int collide(const Body &A, const Body &B, Contact *contacts, double basis[9], double skin = 0.0,
            double *mids = nullptr);

// 6x6 LDL^T solve (maths.js:660-725).
void solve6(const double aLin[9], const double aAng[9], const double aCross[9],
            const double bLin[3], const double bAng[3], double xLin[3], double xAng[3]);

}  // namespace orc
