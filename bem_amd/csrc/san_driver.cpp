// Exercise of libbem_hip's host runtime through its C ABI, for the sanitizer
// builds (make -C bem_amd/csrc san): AddressSanitizer + UndefinedBehavior-
// Sanitizer in one binary, ThreadSanitizer in another, with the host code of
// the library instrumented (the device code is not).
//
// Host part (no device needed): the unknown-order tree, argument checks.
// Device part (when a context can be created): bodies, assembly, the dense,
// mixed-with-refinement and low-rank solves, a sweep, the serving queue fed
// from two threads at once while a third polls, the dipole flux, and
// teardown with requests still queued.  Results are checked for consistency
// (a served or swept solve equals the same request solved alone, bit for
// bit), so a sanitizer run also shows the path still computes.
//
// Exit status 0 when every check passes; the sanitizers abort on their own
// findings.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/bem_hip.h"

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#include <sanitizer/lsan_interface.h>
#define BEM_SAN_LEAK_CHECK 1
#endif
#endif

namespace {

int failures = 0;
#define EXPECT(c)                                                       \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      failures++;                                                       \
    }                                                                   \
  } while (0)

struct Mesh {
  std::vector<double> xyz;
  std::vector<int32_t> tri;
  int nv() const { return (int)(xyz.size() / 3); }
  int nt() const { return (int)(tri.size() / 3); }
};

// Icosahedron subdivided `level` times, vertices on the unit sphere.
Mesh icosphere(int level) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  Mesh m;
  const double v0[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                            {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  const int f0[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                         {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                         {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  auto push = [&](double x, double y, double z) {
    const double r = std::sqrt(x * x + y * y + z * z);
    m.xyz.push_back(x / r);
    m.xyz.push_back(y / r);
    m.xyz.push_back(z / r);
    return m.nv() - 1;
  };
  for (auto &v : v0) push(v[0], v[1], v[2]);
  for (auto &f : f0) m.tri.insert(m.tri.end(), f, f + 3);
  for (int l = 0; l < level; l++) {
    std::map<std::pair<int, int>, int> mid;
    auto middle = [&](int a, int b) {
      auto k = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(k);
      if (it != mid.end()) return it->second;
      const double *p = &m.xyz[3 * a], *q = &m.xyz[3 * b];
      int i = push(p[0] + q[0], p[1] + q[1], p[2] + q[2]);
      mid[k] = i;
      return i;
    };
    std::vector<int32_t> next;
    for (int f = 0; f < m.nt(); f++) {
      int a = m.tri[3 * f], b = m.tri[3 * f + 1], c = m.tri[3 * f + 2];
      int ab = middle(a, b), bc = middle(b, c), ca = middle(c, a);
      const int q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      next.insert(next.end(), q, q + 12);
    }
    m.tri.swap(next);
  }
  return m;
}

// Octahedron flattened to half-height hz and subdivided `level` times
// (midpoints, no projection): its equator is a closed loop of edges sharper
// than 2 rad when hz is small.
Mesh flat_octahedron(int level, double hz) {
  Mesh m;
  const double v0[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, hz}, {0, 0, -hz}};
  const int f0[8][3] = {{0, 2, 4}, {2, 1, 4}, {1, 3, 4}, {3, 0, 4}, {2, 0, 5}, {1, 2, 5}, {3, 1, 5}, {0, 3, 5}};
  for (auto &v : v0) m.xyz.insert(m.xyz.end(), v, v + 3);
  for (auto &f : f0) m.tri.insert(m.tri.end(), f, f + 3);
  for (int l = 0; l < level; l++) {
    std::map<std::pair<int, int>, int> mid;
    auto middle = [&](int a, int b) {
      auto k = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(k);
      if (it != mid.end()) return it->second;
      for (int c = 0; c < 3; c++) m.xyz.push_back(0.5 * (m.xyz[3 * a + c] + m.xyz[3 * b + c]));
      return mid[k] = m.nv() - 1;
    };
    std::vector<int32_t> next;
    for (int f = 0; f < m.nt(); f++) {
      int a = m.tri[3 * f], b = m.tri[3 * f + 1], c = m.tri[3 * f + 2];
      int ab = middle(a, b), bc = middle(b, c), ca = middle(c, a);
      const int q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      next.insert(next.end(), q, q + 12);
    }
    m.tri.swap(next);
  }
  return m;
}

void host_checks() {
  {
    // a closed loop of sharp edges: the tree holds only the part with
    // triangle 0 (the reference's mesher keeps one segment)
    Mesh o = flat_octahedron(3, 0.2);
    const int ne = 3 * o.nt() / 2;
    std::vector<int32_t> order(ne), leaves(ne);
    int n = 0, nl = 0;
    EXPECT(bem_rwg_tree(o.xyz.data(), o.nv(), o.tri.data(), o.nt(), 32, 0.0, order.data(), &n, leaves.data(), &nl) ==
           0);
    EXPECT(n > 0 && n < ne && nl >= 1);
    std::printf("host checks: flattened octahedron, tree of %d of %d RWGs in %d leaves\n", n, ne, nl);
  }
  Mesh s = icosphere(2);
  const int ne = 3 * s.nt() / 2;
  std::vector<int32_t> order(ne), leaves(ne);
  int n = 0, nl = 0;
  EXPECT(bem_rwg_tree(s.xyz.data(), s.nv(), s.tri.data(), s.nt(), 32, 0.0, order.data(), &n, leaves.data(), &nl) == 0);
  EXPECT(n == ne && nl > 1);
  std::vector<char> seen(ne, 0);
  for (int i = 0; i < n; i++) seen[order[i]]++;
  for (int e = 0; e < ne; e++) EXPECT(seen[e] == 1);
  // an open surface and bad arguments are refused
  EXPECT(bem_rwg_tree(s.xyz.data(), s.nv(), s.tri.data(), s.nt() - 1, 32, 0.0, order.data(), &n, leaves.data(),
                      &nl) == BEM_E_MESH);
  EXPECT(bem_rwg_tree(nullptr, s.nv(), s.tri.data(), s.nt(), 32, 0.0, order.data(), &n, leaves.data(), &nl) ==
         BEM_E_INVALID);
  std::vector<int32_t> bad(s.tri);
  bad[5] = s.nv() + 3;
  EXPECT(bem_rwg_tree(s.xyz.data(), s.nv(), bad.data(), s.nt(), 32, 0.0, order.data(), &n, leaves.data(), &nl) ==
         BEM_E_INVALID);
  std::printf("host checks: tree of %d RWGs in %d leaves\n", n, nl);
}

// A Lorentz medium near the reference's SiC parameters.
void lorentz(double eV, double *eps) {
  const double einf = 6.7, wl = 0.120, wt = 0.0983, g = 0.00059;
  const double re = wl * wl - eV * eV, im = -g * eV;
  const double dr = wt * wt - eV * eV, di = -g * eV;
  const double den = dr * dr + di * di;
  eps[0] = einf * (re * dr + im * di) / den;
  eps[1] = einf * (im * dr - re * di) / den;
}

bem_request request(double eV, double gap, int precision = BEM_PREC_FP64, int flags = 0) {
  bem_request q;
  std::memset(&q, 0, sizeof q);
  q.eV = eV;
  q.a = 1e-6;
  q.T1 = 300.0;
  q.T2 = 0.0;
  q.gap[0] = gap;
  lorentz(eV, q.eps1);
  lorentz(eV, q.eps2);
  q.mu1[0] = q.mu2[0] = 1.0;
  q.precision = precision;
  q.flags = flags;
  return q;
}

void device_checks(bem_ctx *ctx) {
  Mesh s = icosphere(2);
  bem_body *b1 = nullptr, *b2 = nullptr;
  EXPECT(bem_body_upload(ctx, s.xyz.data(), s.nv(), s.tri.data(), s.nt(), &b1) == 0);
  EXPECT(bem_body_upload(ctx, s.xyz.data(), s.nv(), s.tri.data(), s.nt(), &b2) == 0);
  if (!b1 || !b2) return;
  int ne = 0;
  EXPECT(bem_body_edges(b1, nullptr, &ne) == 0 && ne == 480);

  // assembly of the pair into a host matrix
  const int M = 4 * ne;
  std::vector<double> Z(2 * (size_t)M * M);
  bem_body *bb[2] = {b1, b2};
  const double shifts[6] = {3.0, 0, 0, 0, 0, 0}, e1[2] = {1, 0}, m1[2] = {1, 0};
  EXPECT(bem_assemble(ctx, 2, bb, shifts, e1, m1, 0.5, Z.data(), M, 0) == 0);
  int64_t hist[6];
  EXPECT(bem_pair_histogram(ctx, 2, bb, shifts, hist) == 0);

  // dense, mixed + refinement, low-rank
  bem_result dense, mixed, low;
  bem_request q = request(0.11, 3.0);
  EXPECT(bem_spectral_power(ctx, b1, b2, &q, &dense) == 0 && dense.status == 0);
  bem_request qm = request(0.11, 3.0, BEM_PREC_MIXED, BEM_SOLVE_REFINE);
  EXPECT(bem_spectral_power(ctx, b1, b2, &qm, &mixed) == 0 && mixed.status == 0);
  EXPECT(std::fabs(mixed.ReS - dense.ReS) <= 1e-8 * std::fabs(dense.ReS));
  bem_request ql = request(0.11, 5.0, BEM_PREC_FP64, BEM_SOLVE_LOWRANK);
  ql.rank_tol = 1e-10;
  EXPECT(bem_spectral_power(ctx, b1, b2, &ql, &low) == 0 && low.status == 0);
  std::printf("solves: dense %.6e, mixed %.6e (%d steps), low-rank rank %d\n", dense.ReS, mixed.ReS, mixed.iters,
              low.rank);

  // a sweep equals the single solves
  std::vector<bem_request> reqs;
  for (int i = 0; i < 4; i++) reqs.push_back(request(0.09 + 0.01 * i, 3.0));
  std::vector<bem_result> swept(reqs.size());
  EXPECT(bem_sweep(ctx, b1, b2, reqs.data(), (int)reqs.size(), swept.data()) == 0);
  bem_result one;
  EXPECT(bem_spectral_power(ctx, b1, b2, &reqs[2], &one) == 0);
  EXPECT(swept[2].ReS == one.ReS);

  // serving: two submitting threads and a polling thread at once (tickets
  // start at 1; the poller waits for each to be issued)
  const int per = 6;
  std::vector<std::atomic<uint64_t>> tickets(2 * per);
  for (auto &t : tickets) t.store(0);
  std::vector<bem_request> sq(2 * per);
  for (int i = 0; i < 2 * per; i++) sq[i] = request(0.085 + 0.005 * i, 3.0 + 0.1 * (i % 3));
  std::atomic<int> submitted(0), polled(0);
  std::vector<bem_result> served(2 * per);
  auto submit = [&](int t0) {
    for (int i = t0; i < 2 * per; i += 2) {
      uint64_t t = 0;
      if (bem_submit(ctx, b1, b2, &sq[i], &t) != 0 || t == 0) {
        std::fprintf(stderr, "submit %d failed\n", i);
        std::exit(3);
      }
      tickets[i].store(t);
      submitted++;
    }
  };
  std::thread poller([&] {
    for (int i = 0; i < 2 * per; i++) {
      uint64_t t;
      while ((t = tickets[i].load()) == 0) std::this_thread::yield();
      int rc;
      do rc = bem_poll(ctx, t, &served[i], 100);
      while (rc == BEM_PENDING);
      if (rc == 0 && served[i].status == 0) polled++;
    }
  });
  std::thread s0(submit, 0), s1(submit, 1);
  int queued = 0, running = 0;
  EXPECT(bem_queue_depth(ctx, &queued, &running) == 0);
  s0.join();
  s1.join();
  poller.join();
  EXPECT(polled.load() == 2 * per);
  bem_result alone;
  EXPECT(bem_spectral_power(ctx, b1, b2, &sq[7], &alone) == 0);
  EXPECT(served[7].ReS == alone.ReS);
  EXPECT(bem_poll(ctx, 987654321u, &alone, 0) == BEM_E_INVALID);
  std::printf("served %d requests (%d queued, %d running at the first check)\n", polled.load(), queued, running);

  // dipole flux
  bem_dipole d[2];
  std::memset(d, 0, sizeof d);
  d[0].p[0] = 1.5;
  d[0].p[2] = 1.2;
  d[0].j[2] = 1.0;
  d[0].region = -1;
  d[1].p[0] = 3.05;
  d[1].j[0] = 1.0;
  d[1].region = 0;
  double flux[4];
  bem_request qd = request(0.1, 3.0);
  EXPECT(bem_dipole_flux(ctx, b1, b2, &qd, 2, d, flux) == 0);

  // requests left queued at teardown
  for (int i = 0; i < 3; i++) {
    uint64_t t;
    EXPECT(bem_submit(ctx, b1, b2, &sq[i], &t) == 0);
  }
  bem_body_free(b1);
  bem_body_free(b2);
}

}  // namespace

int main() {
  std::setvbuf(stdout, nullptr, _IOLBF, 0);   // each line out before a sanitizer report
  host_checks();
  bem_ctx *ctx = nullptr;
  bem_cfg cfg;
  std::memset(&cfg, 0, sizeof cfg);
  cfg.lanes = 2;
  const int rc = bem_ctx_create(&ctx, &cfg);
  if (rc == 0) {
    device_checks(ctx);
    bem_ctx_destroy(ctx);
  } else {
    std::printf("no device (bem_ctx_create: %d): host checks only\n", rc);
  }
  std::printf("%s (%d failed checks)\n", failures ? "FAILED" : "ok", failures);
  std::fflush(nullptr);
#ifdef BEM_SAN_LEAK_CHECK
  __lsan_do_leak_check();   // the at-exit leak check, which _Exit skips (honours detect_leaks)
#endif
  // The contexts are gone.  Leave without the process's finalisers: in them
  // the HIP runtime unloads its device allocator, after which
  // AddressSanitizer's quarantine may still release device blocks the
  // library freed earlier (the self-block arena's 2 GiB segments exceed the
  // quarantine's size) and stops on an internal CHECK.
  std::_Exit(failures ? 1 : 0);
}
