// transfer_farm: the reference's frequency/gap farm (src/transfer-mpi.cpp)
// served by one device context instead of MPI servants.
//
// Inputs in the working directory, as the reference reads them
// (src/transfer-mpi.cpp:208-224, src/FileUtils.h:7-20): a.txt (scales, m per
// geometry unit), gaps.txt (body-1 translations "x y z"), freqs.txt (photon
// energies, eV), temps.txt (pairs "T1 T2", K), whitespace-separated; the
// bodies body1.x3d and body2.x3d with their <Dielectric> media.
//
// Jobs are formed in the reference's order (temperatures, scales, gaps,
// frequencies; :117-133) and all submitted to the serving engine, which keeps
// as many on the device as lanes and memory allow (continuous batching).
// Each result is appended to ./output when it completes, flushed, as one line
// "a f T1 T2 gap P" in the reference's format "%.15g x5 %.14g" (:171), with
// the columns of its recorded output (examples/output) and gap the
// translation's x component.
//
// Restart: lines already in ./output are read first and the jobs whose
// first five columns they hold are not run again, so an interrupted sweep
// resumes where it stopped.  ./output records only the x component of a gap,
// so a restart is refused when gaps.txt holds distinct gaps with equal x
// (their lines could not be told apart).
//
// usage: transfer_farm [fp64|mixed]
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <set>
#include <string>
#include <vector>

#include "bem_compat.h"
#include "x3d.h"

namespace {

template <int N>
std::vector<std::vector<double>> read_tuples(const char *path) {
  std::ifstream f(path);
  if (!f) {
    std::fprintf(stderr, "transfer_farm: cannot open %s\n", path);
    std::exit(2);
  }
  std::vector<std::vector<double>> out;
  std::vector<double> cur;
  std::string tok;
  while (f >> tok) {
    char *e;
    double v = std::strtod(tok.c_str(), &e);
    if (*e) {
      std::fprintf(stderr, "transfer_farm: %s: not a number '%s'\n", path, tok.c_str());
      std::exit(2);
    }
    cur.push_back(v);
    if ((int)cur.size() == N) {
      out.push_back(cur);
      cur.clear();
    }
  }
  return out;
}

std::string key(double a, double f, double T1, double T2, double g) {
  char buf[160];
  std::snprintf(buf, sizeof buf, "%.15g %.15g %.15g %.15g %.15g", a, f, T1, T2, g);
  return buf;
}

struct Job {
  double a, f, T1, T2, gap[3];
  uint64_t ticket;
};

const char *status_name(int st) {
  switch (st) {
    case BEM_E_INVALID: return "invalid argument";
    case BEM_E_DEVICE: return "device error";
    case BEM_E_NOMEM: return "out of device memory or over the HBM budget";
    case BEM_E_SINGULAR: return "singular system";
    case BEM_E_MESH: return "mesh rejected";
    case BEM_E_UNSUPPORTED: return "unsupported size or mode";
    default: return "error";
  }
}

}  // namespace

int main(int argc, char **argv) {
  int precision = BEM_PREC_FP64;
  if (argc > 2 || (argc == 2 && std::strcmp(argv[1], "fp64") && std::strcmp(argv[1], "mixed"))) {
    std::fprintf(stderr, "usage: transfer_farm [fp64|mixed]\n");
    return 2;
  }
  if (argc == 2 && !std::strcmp(argv[1], "mixed")) precision = BEM_PREC_MIXED;
  try {
    auto as = read_tuples<1>("a.txt");
    auto gaps = read_tuples<3>("gaps.txt");
    auto freqs = read_tuples<1>("freqs.txt");
    auto temps = read_tuples<2>("temps.txt");

    X3DShape b1, b2;
    read_x3d("body1.x3d", b1);
    read_x3d("body2.x3d", b2);
    if (!b1.dielectric || !b2.dielectric) {
      std::fprintf(stderr, "transfer_farm: body1.x3d and body2.x3d need a <Dielectric> element\n");
      return 2;
    }
    Params p1, p2;
    p1.params = b1.params;
    p2.params = b2.params;
    Dielectric *d1 = Dielectrics::getDielectric(b1.material, p1);
    Dielectric *d2 = Dielectrics::getDielectric(b2.material, p2);

    std::set<std::string> done;
    {
      std::ifstream prev("output");
      std::string line;
      while (std::getline(prev, line)) {
        double v[6];
        if (std::sscanf(line.c_str(), "%lf %lf %lf %lf %lf %lf", v, v + 1, v + 2, v + 3, v + 4, v + 5) == 6)
          done.insert(key(v[0], v[1], v[2], v[3], v[4]));
      }
    }
    if (!done.empty())
      for (size_t i = 0; i < gaps.size(); i++)
        for (size_t k = i + 1; k < gaps.size(); k++)
          if (gaps[i][0] == gaps[k][0] && (gaps[i][1] != gaps[k][1] || gaps[i][2] != gaps[k][2])) {
            std::fprintf(stderr,
                         "transfer_farm: gaps.txt holds distinct gaps with the same x (%g %g %g and %g %g %g); "
                         "./output records only x, so a restart cannot tell their jobs apart: move ./output away "
                         "to run all jobs\n",
                         gaps[i][0], gaps[i][1], gaps[i][2], gaps[k][0], gaps[k][1], gaps[k][2]);
            return 2;
          }
    std::vector<Job> jobs;
    for (auto &T : temps)
      for (auto &a : as)
        for (auto &g : gaps)
          for (auto &f : freqs)
            if (!done.count(key(a[0], f[0], T[0], T[1], g[0])))
              jobs.push_back(Job{a[0], f[0], T[0], T[1], {g[0], g[1], g[2]}, 0});
    std::fprintf(stderr, "transfer_farm: %zu jobs (%zu already in ./output)\n", jobs.size(), done.size());

    bem_ctx *ctx = bem::context();
    bem_body *body1 = nullptr, *body2 = nullptr;
    bem::check(bem_body_upload(ctx, b1.coord.data(), (int)(b1.coord.size() / 3), b1.face.data(),
                               (int)(b1.face.size() / 3), &body1));
    bem::check(bem_body_upload(ctx, b2.coord.data(), (int)(b2.coord.size() / 3), b2.face.data(),
                               (int)(b2.face.size() / 3), &body2));
    for (auto &j : jobs) {
      // media as src/HeatTransfer.cpp:15-21 sets them (temperatures first)
      d1->setTemperature(j.T1);
      d2->setTemperature(j.T2);
      const Complex e1 = d1->eps(j.f), m1 = d1->mu(j.f), e2 = d2->eps(j.f), m2 = d2->mu(j.f);
      bem_request q = {};
      q.eV = j.f;
      q.a = j.a;
      q.T1 = j.T1;
      q.T2 = j.T2;
      for (int c = 0; c < 3; c++) q.gap[c] = j.gap[c];
      q.eps1[0] = e1.real(), q.eps1[1] = e1.imag(), q.mu1[0] = m1.real(), q.mu1[1] = m1.imag();
      q.eps2[0] = e2.real(), q.eps2[1] = e2.imag(), q.mu2[0] = m2.real(), q.mu2[1] = m2.imag();
      q.precision = precision;
      if (bem_submit(ctx, body1, body2, &q, &j.ticket) != 0) {   // e.g. larger than the device: this job only
        std::fprintf(stderr, "transfer_farm: job a=%g f=%g gap=%g not submitted: %s\n", j.a, j.f, j.gap[0],
                     bem_last_error(ctx));
        j.ticket = 0;
      }
    }
    FILE *out = std::fopen("output", "a");
    if (!out) {
      std::fprintf(stderr, "transfer_farm: cannot append to ./output\n");
      return 2;
    }
    // completion order: every outstanding ticket is polled, waiting briefly
    // only on the oldest
    std::vector<bool> finished(jobs.size(), false);
    size_t left = jobs.size();
    int failures = 0;
    for (size_t i = 0; i < jobs.size(); i++)
      if (jobs[i].ticket == 0) {
        finished[i] = true;
        left--;
        failures++;
      }
    while (left) {
      bool first = true;
      for (size_t i = 0; i < jobs.size(); i++) {
        if (finished[i]) continue;
        bem_result r;
        const int rc = bem_poll(ctx, jobs[i].ticket, &r, first ? 20 : 0);
        first = false;
        if (rc == BEM_PENDING) continue;
        finished[i] = true;
        left--;
        if (rc != 0 || r.status != 0) {
          const int st = r.status ? r.status : rc;   // the ticket's own status, not the context's last message
          std::fprintf(stderr, "transfer_farm: job a=%g f=%g gap=%g %g %g failed: status %d (%s)\n", jobs[i].a,
                       jobs[i].f, jobs[i].gap[0], jobs[i].gap[1], jobs[i].gap[2], st, status_name(st));
          failures++;
          continue;
        }
        std::fprintf(out, "%s %.14g\n", key(jobs[i].a, jobs[i].f, jobs[i].T1, jobs[i].T2, jobs[i].gap[0]).c_str(),
                     r.P);
        std::fflush(out);
      }
    }
    std::fclose(out);
    bem_body_free(body1);
    bem_body_free(body2);
    return failures ? 1 : 0;
  } catch (const bem::Error &e) {
    std::fprintf(stderr, "transfer_farm: %s\n", e.what());
    return 1;
  }
}
