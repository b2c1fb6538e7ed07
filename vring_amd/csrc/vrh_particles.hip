// vrh_particles.hip -- Lagrangian particle step (part_tools.F90:374-1025).
#include <hip/hip_runtime.h>

#include "vrh_internal.h"

void launch_particles(vrh_t *h, int irk) { (void)h; (void)irk; }

extern "C" {
int vrh_particles_create(vrh_t *h, int, const long *, const double *, const int *, int) {
  if (h) h->err = "particles: not implemented yet";
  return -1;
}
int vrh_particles_upload(vrh_t *h, int, long, const double *, const int *) { return h ? -1 : -1; }
int vrh_particles_download(vrh_t *h, int, long, double *, int *) { return h ? -1 : -1; }
int vrh_particles_step(vrh_t *h, int) { return h ? -1 : -1; }
}
