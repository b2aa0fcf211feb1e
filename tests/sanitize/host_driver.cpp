// Drives the host-only entry points of libnekrs_hip (the .re2 parser, the
// unstructured numbering, the midside map, the multigrid schedule parser, the
// FDM setup helpers, the coarse-id and GLL helpers) on the files and strings
// given, for a build with AddressSanitizer and UndefinedBehaviorSanitizer on
// the host code.  Error codes are expected for malformed input; a sanitizer
// report (or a crash) is the failure.  No GPU call is made.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nrs_hip.h"

static int run_re2(const char *path)
{
  int nelgt = 0, nelgv = 0, ncurve = 0, nbc = 0;
  int rc = nrs_re2_info(path, &nelgt, &nelgv, &ncurve, &nbc);
  std::printf("%s: info rc %d (E %d, curves %d, bc %d)\n", path, rc, nelgt, ncurve, nbc);
  if (rc != NRS_OK) return 0;
  std::vector<double> vx(8 * (size_t)nelgt), vy(vx.size()), vz(vx.size());
  std::vector<int> group(nelgt), peer(12 * (size_t)nelgt);
  std::vector<char> cbc(18 * (size_t)nelgt);
  rc = nrs_re2_read(path, vx.data(), vy.data(), vz.data(), group.data(), cbc.data());
  std::printf("  read rc %d\n", rc);
  if (rc != NRS_OK) return 0;
  std::vector<int> elem(ncurve + 1), side(ncurve + 1);
  std::vector<double> params(5 * (size_t)ncurve + 5);
  std::vector<char> type(ncurve + 1);
  rc = nrs_re2_curves(path, elem.data(), side.data(), params.data(), type.data());
  std::printf("  curves rc %d\n", rc);
  const int crc = rc;
  rc = nrs_re2_periodic(path, peer.data());
  std::printf("  periodic rc %d\n", rc);
  for (int N : {1, 3}) {
    const size_t L = (size_t)nelgt * (N + 1) * (N + 1) * (N + 1);
    std::vector<double> x(L), y(L), z(L);
    std::vector<long long> ids(L);
    rc = nrs_mesh_from_vertices(N, nelgt, vx.data(), vy.data(), vz.data(), 0.0, x.data(), y.data(), z.data(),
                                ids.data());
    std::printf("  mesh N=%d rc %d\n", N, rc);
    if (rc == NRS_OK && crc == NRS_OK) {
      rc = nrs_mesh_midside(N, nelgt, vx.data(), vy.data(), vz.data(), ncurve, elem.data(), side.data(),
                            params.data(), type.data(), x.data(), y.data(), z.data());
      std::printf("  midside N=%d rc %d\n", N, rc);
    }
    if (rc == NRS_OK) {
      rc = nrs_mesh_from_vertices_periodic(N, nelgt, vx.data(), vy.data(), vz.data(), 0.0, peer.data(),
                                           x.data(), y.data(), z.data(), ids.data());
      std::printf("  periodic mesh N=%d rc %d\n", N, rc);
    }
  }
  return 0;
}

static void run_schedules()
{
  const char *sch[] = {"p=7+degree=2, p=3+degree=2, p=1, p=3+degree=2, p=7+degree=2",
                       "p=7, p=5, p=3, p=1", "p=1+degree=5, p=7", "", "p=", "p=7+degree=",
                       "p=7+degree=99", "p=-3", "p=7,,p=1", "degree=2", "p=7+degree=2+degree=3",
                       "p=99999999999999999999", "p=7, p=7, p=1", "   p = 7 , p = 3 , p = 1  "};
  for (const char *s : sch) {
    nrs_elliptic_opts o;
    nrs_elliptic_default_opts(&o);
    const int rc = nrs_parse_mg_schedule(s, 7, 2, &o);
    std::printf("schedule '%s': rc %d levels %d\n", s, rc, o.nLevels);
  }
}

static void run_fdm_and_helpers()
{
  for (int N = 1; N <= 9; ++N) {
    const int n = N + 3;
    std::vector<double> S((size_t)n * n), lam(n);
    for (int lbc = 0; lbc <= 1; ++lbc)
      for (int rbc = 0; rbc <= 1; ++rbc) {
        const int rc = nrs_fdm_1d(N, lbc, rbc, 0.3, 0.5, 0.2, S.data(), lam.data());
        if (rc != NRS_OK) std::printf("fdm_1d N=%d rc %d\n", N, rc);
      }
    std::vector<double> z(N + 1), w(N + 1), D((size_t)(N + 1) * (N + 1));
    nrs_gll(N, z.data(), w.data(), D.data());
    const int E = 8, Np = (N + 1) * (N + 1) * (N + 1);
    std::vector<double> x((size_t)E * Np), y(x.size()), zc(x.size());
    std::vector<long long> ids(x.size());
    std::vector<int> etob(E * 6);
    nrs_box_mesh(2, 2, 2, N, 2, 0.3, x.data(), y.data(), zc.data(), ids.data(), etob.data());
    std::vector<double> len((size_t)9 * E);
    const int lrc = nrs_fdm_element_lengths(N, E, x.data(), y.data(), zc.data(), ids.data(), len.data());
    if (lrc != NRS_OK) std::printf("element lengths N=%d rc %d\n", N, lrc);
    for (int Nc = 1; Nc < N; ++Nc) {
      std::vector<long long> idc((size_t)E * (Nc + 1) * (Nc + 1) * (Nc + 1));
      const int rc = nrs_coarse_global_ids(N, E, ids.data(), Nc, idc.data());
      if (rc != NRS_OK) std::printf("coarse ids %d->%d rc %d\n", N, Nc, rc);
    }
  }
  std::printf("fdm and helpers done\n");
}

int main(int argc, char **argv)
{
  for (int a = 1; a < argc; ++a) run_re2(argv[a]);
  run_schedules();
  run_fdm_and_helpers();
  std::printf("host driver done\n");
  return 0;
}
