/* Shared helpers for the fixture drivers that run this project's test
 * problems through the reference's own sequential OPS library (libops_seq).
 * Test infrastructure only.  Outputs are raw little-endian arrays written to
 * a directory; tests/golden/make_golden.py packs them into fixtures. */
#ifndef REFDRV_COMMON_H
#define REFDRV_COMMON_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>

static std::string g_outdir = ".";

static void drv_write(const std::string &name, const void *p, size_t bytes) {
  std::string path = g_outdir + "/" + name + ".bin";
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) { perror(path.c_str()); exit(2); }
  if (bytes && fwrite(p, 1, bytes, f) != bytes) { perror("fwrite"); exit(2); }
  fclose(f);
}

template <class T>
static void drv_write_vec(const std::string &name, const std::vector<T> &v) {
  drv_write(name, v.data(), v.size() * sizeof(T));
}

/* Interior fetch through the reference API (ops_dat_fetch_data copies the
 * interior only: ops/c/src/core/ops_dummy_singlenode.cpp:741-765). */
template <class T>
static std::vector<T> drv_fetch(ops_dat d, size_t npts) {
  std::vector<T> v(npts);
  ops_dat_fetch_data(d, 0, (char *)v.data());
  return v;
}

#endif
