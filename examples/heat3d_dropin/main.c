/* A user driver written against the generated C API only (the way the
 * reference's example drivers are): Formura_Init, a Formura_Forward loop,
 * direct reads of formura_data and to_pos_*, Formura_Finalize.  It links
 * unchanged against the reference's generated C or against the HIP library.
 * Prints a checksum of the final state and the Navi offsets. */
#include <stdio.h>
#include <stdint.h>
#include "heat3d_64.h"

#define T_MAX 40

int main(int argc, char **argv) {
  Formura_Navi n;
  Formura_Init(&argc, &argv, &n);
  /* overwrite the init graph's gaussian with a deterministic field */
  uint64_t s = 12345;
  for (int ix = n.lower_x; ix < n.upper_x; ix++)
    for (int iy = n.lower_y; iy < n.upper_y; iy++)
      for (int iz = n.lower_z; iz < n.upper_z; iz++) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        formura_data.q[ix][iy][iz] = (double)(s >> 11) * (1.0 / 9007199254740992.0);
      }
  while (n.time_step < T_MAX) Formura_Forward(&n);
  double sum = 0.0, wsum = 0.0;
  for (int ix = n.lower_x; ix < n.upper_x; ix++)
    for (int iy = n.lower_y; iy < n.upper_y; iy++)
      for (int iz = n.lower_z; iz < n.upper_z; iz++) {
        sum += formura_data.q[ix][iy][iz];
        wsum += formura_data.q[ix][iy][iz] * to_pos_x(ix, n);
      }
  printf("time_step %d offset %d %d %d sum %.17g wsum %.17g q000 %a\n", n.time_step, n.offset_x, n.offset_y,
         n.offset_z, sum, wsum, formura_data.q[0][0][0]);
  Formura_Finalize();
  return 0;
}
