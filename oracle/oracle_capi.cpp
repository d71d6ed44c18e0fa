// oracle_capi.cpp — flat C entry points over the f64 oracle, for ctypes.
//
// TEST INFRASTRUCTURE: loaded only by tests/, the smoke check and bench.py's
// CPU-baseline leg. Scene construction mirrors the reference's
// test/parity/reference_main.cpp:45-130 (body / joint records) and
// src/physics/rigid.js (box and sphere constructors).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "avbd_oracle.hpp"

using namespace orc;

extern "C" {

void *orc_create(void) { return new World(); }
void orc_destroy(void *w) { delete static_cast<World *>(w); }

// d: dt, gravity, betaLin, betaAng, alpha, gamma, penaltyMin, penaltyMax,
//    collisionMargin, stickThreshold
// i: iterations, postStabilize, rotatedInertia, cachedContactJacobians
void orc_set_params(void *w, const double *d, const int *i) {
  Params &P = static_cast<World *>(w)->prm;
  P.dt = d[0]; P.gravity = d[1]; P.betaLin = d[2]; P.betaAng = d[3]; P.alpha = d[4];
  P.gamma = d[5]; P.penaltyMin = d[6]; P.penaltyMax = d[7]; P.collisionMargin = d[8];
  P.stickThreshold = d[9];
  P.iterations = i[0]; P.postStabilize = i[1] != 0; P.rotatedInertia = i[2] != 0;
  P.cachedContactJacobians = i[3] != 0;
}

// The reference GPU backend's contact model (include/avbd.h
// AVBD_CONTACTS_GPU): model 1 switches it on, with the solver's contact
// offset, rest offset, restitution threshold and persistence distance
// (solver.js:142-162).
void orc_set_contact_model(void *w, int model, double contactOffset, double restOffset, double restitutionThreshold,
                           double contactPersistence) {
  Params &P = static_cast<World *>(w)->prm;
  P.contactModel = model;
  P.contactOffset = contactOffset;
  P.restOffset = restOffset;
  P.restitutionThreshold = restitutionThreshold;
  P.contactPersistence = contactPersistence;
}

// Rigid.setMaterial (rigid.js:144-214) for body i: static and dynamic
// friction, restitution, and the contact and rest offsets (NaN: the
// solver's). Returns -1 for an index out of range and -2 for a material
// setMaterial rejects (a negative or non-finite coefficient, restitution
// outside [0, 1], dynamic above static friction, a rest offset above the
// contact offset when both are given); the body is then unchanged.
int orc_set_material(void *w, int i, double staticFriction, double dynamicFriction, double restitution,
                     double contactOffset, double restOffset) {
  World &W = *static_cast<World *>(w);
  if (i < 0 || i >= int(W.bodies.size())) return -1;
  auto nonNeg = [](double v) { return std::isfinite(v) && v >= 0; };
  if (!nonNeg(staticFriction) || !nonNeg(dynamicFriction) || dynamicFriction > staticFriction) return -2;
  if (!std::isfinite(restitution) || restitution < 0 || restitution > 1) return -2;
  if (!std::isnan(contactOffset) && !nonNeg(contactOffset)) return -2;
  if (!std::isnan(restOffset) && !std::isfinite(restOffset)) return -2;
  if (!std::isnan(contactOffset) && !std::isnan(restOffset) && restOffset > contactOffset) return -2;
  Body &b = W.bodies[i];
  b.staticFriction = staticFriction;
  b.dynamicFriction = dynamicFriction;
  b.restitution = restitution;
  b.contactOffset = contactOffset;
  b.restOffset = restOffset;
  return 0;
}

static void setPose(Body &b, const double *q, const double *w) {
  for (int k = 0; k < 4; k++) b.q[k] = b.q0[k] = q[k];
  for (int k = 0; k < 3; k++) b.w[k] = w[k];
}

int orc_add_box(void *w, const double *size, double density, double friction, const double *pos,
                const double *vel, const double *quat, const double *angvel) {
  World &W = *static_cast<World *>(w);
  const int id = W.addBox(size, density, friction, pos, vel);
  setPose(W.bodies[id], quat, angvel);
  return id;
}

int orc_add_sphere(void *w, double diameter, double density, double friction, const double *pos,
                   const double *vel, const double *quat, const double *angvel) {
  World &W = *static_cast<World *>(w);
  const int id = W.addSphere(diameter, density, friction, pos, vel);
  setPose(W.bodies[id], quat, angvel);
  return id;
}

int orc_add_joint(void *w, int a, int b, const double *rA, const double *rB, double stiffLin, double stiffAng,
                  double fracture) {
  World &W = *static_cast<World *>(w);
  if (b < 0 || b >= int(W.bodies.size()) || a >= int(W.bodies.size())) return -1;
  return W.addJoint(a, b, rA, rB, stiffLin, stiffAng, fracture);
}

int orc_add_spring(void *w, int a, int b, const double *rA, const double *rB, double stiffness, double rest,
                   double tearStrain) {
  World &W = *static_cast<World *>(w);
  const int n = int(W.bodies.size());
  if (a < 0 || b < 0 || a >= n || b >= n) return -1;
  return W.addSpring(a, b, rA, rB, stiffness, rest, tearStrain);
}

// A joint's anchor rA (a world point when it is world-anchored); joint is the
// joint's creation rank among the joints and springs.
int orc_set_joint_anchor(void *w, int joint, const double *rA) {
  World &W = *static_cast<World *>(w);
  int rank = 0;
  for (Force &f : W.pool) {
    if (f.kind == MANIFOLD) continue;
    if (rank++ == joint) {
      if (f.kind != JOINT) return -1;
      for (int k = 0; k < 3; k++) f.jrA[k] = rA[k];
      return 0;
    }
  }
  return -1;
}

// Spring force idx of the force list: penalty, C, torn, peak strain, rest.
int orc_get_spring(void *w, int idx, double *out) {
  const World &W = *static_cast<World *>(w);
  if (idx < 0 || idx >= int(W.forces.size())) return -1;
  const Force &f = W.pool[W.forces[idx]];
  if (f.kind != SPRING) return -1;
  out[0] = f.penalty;
  out[1] = f.springC;
  out[2] = f.broken ? 1.0 : 0.0;
  out[3] = f.peakStrain;
  out[4] = f.rest;
  return 0;
}

void orc_set_schedule(void *w, int mode, const int *order, int n) {
  World &W = *static_cast<World *>(w);
  W.schedule = Schedule(mode);
  W.bodyOrder.assign(order ? order : nullptr, order ? order + n : nullptr);
}

void orc_set_broadphase(void *w, int mode) { static_cast<World *>(w)->broadphase = mode; }
void orc_set_color_order(void *w, int order) { static_cast<World *>(w)->colorOrder = order; }
unsigned orc_color_priority(unsigned i) { return colorPriority(i); }

// Overwrite one body's pose and velocities (a launch or a pick-and-throw).
void orc_set_body(void *w, int i, const double *pos, const double *quat, const double *vel, const double *angvel) {
  Body &b = static_cast<World *>(w)->bodies[i];
  for (int k = 0; k < 3; k++) {
    if (pos) b.p[k] = pos[k];
    if (vel) b.v[k] = b.vprev[k] = vel[k];
    if (angvel) b.w[k] = angvel[k];
  }
  if (quat) for (int k = 0; k < 4; k++) b.q[k] = quat[k];
}
// Removes the bodies whose doomed[i] is nonzero (Solver.removeBodies).
int orc_remove_bodies(void *w, const int *doomed, int n) {
  World &W = *static_cast<World *>(w);
  if (n != int(W.bodies.size())) return -1;
  std::vector<char> d(doomed, doomed + n);
  return W.removeBodies(d);
}
void orc_record_pairs(void *w, int on) { static_cast<World *>(w)->recordPairs = on != 0; }

// Runs n steps; returns the wall time in seconds.
double orc_step(void *w, int n) {
  World &W = *static_cast<World *>(w);
  const auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < n; s++) W.step();
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int orc_num_bodies(void *w) { return int(static_cast<World *>(w)->bodies.size()); }
int orc_num_dynamic(void *w) {
  int n = 0;
  for (const Body &b : static_cast<World *>(w)->bodies) n += b.mass > 0;
  return n;
}

// Reads a `.scene` file into the world (test/parity/reference_main.cpp:45-124,
// the same records as test/parity/scene_loader.mjs:20-84): '#' starts a
// comment, tokens split on blanks, tabs, CR and LF, "inf" and "-inf" are the
// infinities and every other token goes through strtod.
//   body   sx sy sz density friction px py pz vx vy vz qx qy qz qw wx wy wz
//   joint  idxA idxB rax ray raz rbx rby rbz stiffLin stiffAng fracture
//   spring idxA idxB rax ray raz rbx rby rbz stiffness rest
// Returns the number of bodies read; -1 when the file cannot be opened, -2
// for a short record or a joint or spring naming a body not yet read, and -4
// for an unknown record.
static double sceneValue(const char *t) {
  if (std::strcmp(t, "inf") == 0) return std::numeric_limits<double>::infinity();
  if (std::strcmp(t, "-inf") == 0) return -std::numeric_limits<double>::infinity();
  return std::strtod(t, nullptr);
}

int orc_load_scene(void *w, const char *path) {
  World &W = *static_cast<World *>(w);
  FILE *f = std::fopen(path, "r");
  if (!f) return -1;
  std::vector<int> ids;
  char line[8192];
  int rc = 0;
  while (rc == 0 && std::fgets(line, sizeof(line), f)) {
    if (char *h = std::strchr(line, '#')) *h = '\0';
    std::vector<const char *> tok;
    for (char *t = std::strtok(line, " \t\r\n"); t; t = std::strtok(nullptr, " \t\r\n")) tok.push_back(t);
    if (tok.empty()) continue;
    if (std::strcmp(tok[0], "body") == 0) {
      if (tok.size() < 19) { rc = -2; break; }
      double v[18];
      for (int i = 0; i < 18; i++) v[i] = sceneValue(tok[1 + i]);
      const int id = W.addBox(v, v[3], v[4], v + 5, v + 8);
      setPose(W.bodies[id], v + 11, v + 15);
      ids.push_back(id);
    } else if (std::strcmp(tok[0], "joint") == 0) {
      if (tok.size() < 12) { rc = -2; break; }
      const int ia = int(sceneValue(tok[1])), ib = int(sceneValue(tok[2]));
      if (ib < 0 || ib >= int(ids.size()) || ia >= int(ids.size())) { rc = -2; break; }
      double v[9];
      for (int i = 0; i < 9; i++) v[i] = sceneValue(tok[3 + i]);
      W.addJoint(ia < 0 ? -1 : ids[ia], ids[ib], v, v + 3, v[6], v[7], v[8]);
    } else if (std::strcmp(tok[0], "spring") == 0) {
      // spring idxA idxB rax ray raz rbx rby rbz stiffness rest
      if (tok.size() < 11) { rc = -2; break; }
      const int ia = int(sceneValue(tok[1])), ib = int(sceneValue(tok[2]));
      if (ia < 0 || ib < 0 || ib >= int(ids.size()) || ia >= int(ids.size())) { rc = -2; break; }
      double v[8];
      for (int i = 0; i < 8; i++) v[i] = sceneValue(tok[3 + i]);
      W.addSpring(ids[ia], ids[ib], v, v + 3, v[6], v[7], std::numeric_limits<double>::infinity());
    } else {
      rc = -4;
    }
  }
  std::fclose(f);
  return rc == 0 ? int(ids.size()) : rc;
}

// One frame of the trajectory dump of test/parity/reference_main.cpp:126-140:
// a line per body, "step index p[3] q[4] v[3] w[3]", numbers as printf's
// %.17g. Writes up to cap bytes (NUL-terminated when cap > 0) and returns the
// length of the whole frame.
long orc_format_state(void *w, int step, char *buf, long cap) {
  const World &W = *static_cast<World *>(w);
  std::string out;
  char line[512];
  for (size_t i = 0; i < W.bodies.size(); i++) {
    const Body &b = W.bodies[i];
    const int len = std::snprintf(
        line, sizeof(line),
        "%d %zu %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", step, i,
        b.p[0], b.p[1], b.p[2], b.q[0], b.q[1], b.q[2], b.q[3], b.v[0], b.v[1], b.v[2], b.w[0], b.w[1], b.w[2]);
    out.append(line, size_t(len));
  }
  if (cap > 0) {
    const size_t n = std::min(out.size(), size_t(cap - 1));
    std::memcpy(buf, out.data(), n);
    buf[n] = '\0';
  }
  return long(out.size());
}

// Per body, 16 doubles: p[3] q[4] v[3] w[3] vprev[3].
void orc_get_state(void *w, double *out) {
  const World &W = *static_cast<World *>(w);
  for (size_t i = 0; i < W.bodies.size(); i++) {
    const Body &b = W.bodies[i];
    double *o = out + i * 16;
    for (int k = 0; k < 3; k++) o[k] = b.p[k];
    for (int k = 0; k < 4; k++) o[3 + k] = b.q[k];
    for (int k = 0; k < 3; k++) { o[7 + k] = b.v[k]; o[10 + k] = b.w[k]; o[13 + k] = b.vprev[k]; }
  }
}

// Per body, 8 doubles: mass moment[3] radius shape size[0..1]... (mass props
// for tests): mass, moment[3], radius, shape, density, friction.
void orc_get_mass(void *w, double *out) {
  const World &W = *static_cast<World *>(w);
  for (size_t i = 0; i < W.bodies.size(); i++) {
    const Body &b = W.bodies[i];
    double *o = out + i * 8;
    o[0] = b.mass; o[1] = b.moment[0]; o[2] = b.moment[1]; o[3] = b.moment[2];
    o[4] = b.radius; o[5] = b.shape; o[6] = b.density; o[7] = b.friction;
  }
}

int orc_num_forces(void *w) { return int(static_cast<World *>(w)->forces.size()); }

// Per live force in list order: kind, a, b, nc (ints) and seq (int64).
void orc_get_forces(void *w, int *kab, long long *seq) {
  const World &W = *static_cast<World *>(w);
  for (size_t i = 0; i < W.forces.size(); i++) {
    const Force &f = W.pool[W.forces[i]];
    kab[i * 4] = f.kind; kab[i * 4 + 1] = f.a; kab[i * 4 + 2] = f.b;
    kab[i * 4 + 3] = f.kind == MANIFOLD ? f.nc : 0;
    seq[i] = f.seq;
  }
}

// Manifold at list position `idx`: basis[9], mu, and per contact (19 doubles):
// rA[3] rB[3] C0[3] k[3] lam[3] C[3] stick; features separately.
int orc_get_manifold(void *w, int idx, double *basis_mu, unsigned *features, double *contacts) {
  const World &W = *static_cast<World *>(w);
  if (idx < 0 || idx >= int(W.forces.size())) return -1;
  const Force &f = W.pool[W.forces[idx]];
  if (f.kind != MANIFOLD) return -1;
  for (int k = 0; k < 9; k++) basis_mu[k] = f.basis[k];
  basis_mu[9] = f.mu;
  for (int i = 0; i < f.nc; i++) {
    const Contact &c = f.c[i];
    features[i] = c.feature;
    double *o = contacts + i * 19;
    for (int k = 0; k < 3; k++) {
      o[k] = c.rA[k]; o[3 + k] = c.rB[k]; o[6 + k] = c.C0[k];
      o[9 + k] = c.k[k]; o[12 + k] = c.lam[k]; o[15 + k] = c.C[k];
    }
    o[18] = c.stick ? 1.0 : 0.0;
  }
  return f.nc;
}

// The GPU contact model's fields of the manifold at list position `idx`:
// extra = {dynamic friction, contact distance}, bias[c] = each contact's
// restitution target. Returns the contact count.
int orc_get_manifold_gpu(void *w, int idx, double *extra, double *bias) {
  const World &W = *static_cast<World *>(w);
  if (idx < 0 || idx >= int(W.forces.size())) return -1;
  const Force &f = W.pool[W.forces[idx]];
  if (f.kind != MANIFOLD) return -1;
  extra[0] = f.muD;
  extra[1] = f.contactDistance;
  for (int i = 0; i < f.nc; i++) bias[i] = f.c[i].bias;
  return f.nc;
}

long long orc_geometric_matches(void *w) { return static_cast<World *>(w)->geometricMatches; }

// Joint at list position `idx`: C0Lin kLin lamLin CLin C0Ang kAng lamAng CAng (24) + broken.
int orc_get_joint(void *w, int idx, double *out) {
  const World &W = *static_cast<World *>(w);
  if (idx < 0 || idx >= int(W.forces.size())) return -1;
  const Force &f = W.pool[W.forces[idx]];
  if (f.kind != JOINT) return -1;
  const double *src[8] = {f.C0Lin, f.kLin, f.lamLin, f.CLin, f.C0Ang, f.kAng, f.lamAng, f.CAng};
  for (int s = 0; s < 8; s++)
    for (int k = 0; k < 3; k++) out[s * 3 + k] = src[s][k];
  out[24] = f.broken ? 1.0 : 0.0;
  return 0;
}

// Colouring of the live force graph as it stands (coloring.js:43-118).
int orc_coloring(void *w, int *colorOf, int *offsets, int *entries) {
  const World &W = *static_cast<World *>(w);
  Coloring c;
  W.buildColoring(c);
  std::memcpy(colorOf, c.colorOf.data(), sizeof(int) * c.colorOf.size());
  if (offsets) std::memcpy(offsets, c.offsets.data(), sizeof(int) * c.offsets.size());
  if (entries) std::memcpy(entries, c.entries.data(), sizeof(int) * c.entries.size());
  return c.numColors;
}

// Colouring used by the last step under the colour schedule.
int orc_last_coloring(void *w, int *colorOf) {
  const World &W = *static_cast<World *>(w);
  std::memcpy(colorOf, W.coloring.colorOf.data(), sizeof(int) * W.coloring.colorOf.size());
  return W.coloring.numColors;
}

int orc_num_pairs(void *w) { return int(static_cast<World *>(w)->lastPairs.size() / 2); }
void orc_get_pairs(void *w, int *out) {
  const World &W = *static_cast<World *>(w);
  std::memcpy(out, W.lastPairs.data(), sizeof(int) * W.lastPairs.size());
}

// Narrow phase between bodies a and b of the world (collide.js:612-685).
// contacts: 7 per contact (feature as double, rA[3], rB[3]).
int orc_collide(void *w, int a, int b, double *basis, double *contacts) {
  const World &W = *static_cast<World *>(w);
  Contact c[MAX_CONTACTS];
  for (auto &x : c) x.reset();
  const int n = collide(W.bodies[a], W.bodies[b], c, basis);
  for (int i = 0; i < n; i++) {
    contacts[i * 7] = double(c[i].feature);
    for (int k = 0; k < 3; k++) { contacts[i * 7 + 1 + k] = c[i].rA[k]; contacts[i * 7 + 4 + k] = c[i].rB[k]; }
  }
  return n;
}

void orc_solve6(const double *aLin, const double *aAng, const double *aCross, const double *bLin,
                const double *bAng, double *xLin, double *xAng) {
  solve6(aLin, aAng, aCross, bLin, bAng, xLin, xAng);
}

}  // extern "C"
