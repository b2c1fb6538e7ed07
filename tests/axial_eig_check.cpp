// Host-only harness for the axial-mesh pressure transforms of libvrh
// (host_axial_transforms, an internal C++ symbol of the library): reads nz
// and the z differences m3, c3 (k = 0..nz+2) from stdin, writes lam, V^-1 and
// V (row-major) to stdout, one value per line.  Makes no GPU call.
#include <cstdio>
#include <string>
#include <vector>

int host_axial_transforms(const double *m3, const double *c3, int nz, std::vector<double> &vi,
                          std::vector<double> &vf, std::vector<double> &lam, std::string &err);

int main() {
  int nz = 0;
  if (scanf("%d", &nz) != 1 || nz < 2) return 2;
  std::vector<double> m3(nz + 3), c3(nz + 3);
  for (auto &x : m3) if (scanf("%lf", &x) != 1) return 2;
  for (auto &x : c3) if (scanf("%lf", &x) != 1) return 2;
  std::vector<double> vi, vf, lam;
  std::string err;
  if (host_axial_transforms(m3.data(), c3.data(), nz, vi, vf, lam, err)) {
    fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  for (double x : lam) printf("%.17e\n", x);
  for (double x : vi) printf("%.17e\n", x);
  for (double x : vf) printf("%.17e\n", x);
  return 0;
}
