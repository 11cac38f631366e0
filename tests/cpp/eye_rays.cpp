// Writes the eye rays of srf::EyeRays (include/srf_frame.hpp) for one camera:
//   eye_rays ex ey ez lx ly lz ux uy uz fov width height out.bin
// out.bin: width*height records of (org[3], dir[3]) floats, then one valid
// byte per pixel.
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "srf_frame.hpp"

int main(int argc, char **argv) {
  if (argc != 14) return 2;
  srf::Camera c;
  for (int k = 0; k < 3; k++) {
    c.eye[k] = strtof(argv[1 + k], nullptr);
    c.lookat[k] = strtof(argv[4 + k], nullptr);
    c.up[k] = strtof(argv[7 + k], nullptr);
  }
  c.fov = strtof(argv[10], nullptr);
  const int w = atoi(argv[11]), h = atoi(argv[12]);
  std::vector<srf_ray> rays((size_t)w * h);
  std::vector<uint8_t> valid((size_t)w * h);
  srf::EyeRays(c, w, h, rays.data(), valid.data());
  FILE *fp = fopen(argv[13], "wb");
  if (!fp) return 1;
  for (const srf_ray &r : rays) {
    fwrite(r.org, 4, 3, fp);
    fwrite(r.dir, 4, 3, fp);
  }
  fwrite(valid.data(), 1, valid.size(), fp);
  fclose(fp);
  return 0;
}
