/* Halo-transfer fixture driver: runs ops_halo_transfer of the reference's
 * sequential OPS library (ops/c/src/sequential/ops_host_singlenode.cpp:
 * 172-446) on cases read from a text spec, and dumps every dat's whole
 * array (halos included, x padding excluded) before the first group and
 * after each group.  Test infrastructure only.
 *
 * Spec lines (whitespace separated):
 *   case NAME NDIM
 *   dat NAME TYPE DIM s0 s1 s2 dm0 dm1 dm2 dp0 dp1 dp2     (TYPE double|float)
 *   group
 *   halo FROM TO it0 it1 it2 fb0 fb1 fb2 tb0 tb1 tb2 fd0 fd1 fd2 td0 td1 td2
 *   end
 * Entries beyond NDIM are ignored.  Dat contents before the first group are
 * uniform values of oracle/rng.h (seed 7, stream = 100 * case + dat, index =
 * raw position x fastest, component innermost), rounded to float for float
 * dats.  Outputs: halo_<case>_<dat>_g<k>.bin, k = 0 before any group.
 *
 * usage: ref_halo SPEC OUTDIR
 */
#define OPS_3D
#include <ops_seq_v2.h>

#include <fstream>
#include <map>
#include <sstream>

#include "refdrv_common.h"
#include "rng.h"

namespace {

struct DatSpec {
  std::string name, type;
  int dim, size[3], dm[3], dp[3];
  ops_dat dat;
};

// extent of the raw array without the x padding, per dimension
void ext_of(const DatSpec &s, int nd, int e[3]) {
  for (int i = 0; i < 3; i++) e[i] = i < nd ? s.size[i] - s.dm[i] + s.dp[i] : 1;
}

char *raw_at(ops_dat d, int x, int y, int z, int c) {
  const size_t pos = (size_t)x + (size_t)d->size[0] * ((size_t)y + (size_t)d->size[1] * (size_t)z);
  return d->data + pos * d->elem_size + (size_t)c * d->type_size;
}

void fill(const DatSpec &s, int nd, uint64_t stream) {
  int e[3];
  ext_of(s, nd, e);
  uint64_t i = 0;
  for (int z = 0; z < e[2]; z++)
    for (int y = 0; y < e[1]; y++)
      for (int x = 0; x < e[0]; x++)
        for (int c = 0; c < s.dim; c++, i++) {
          const double v = 2.0 * orc_rng_u01(7, stream, i) - 1.0;
          if (s.type == "float") *(float *)raw_at(s.dat, x, y, z, c) = (float)v;
          else *(double *)raw_at(s.dat, x, y, z, c) = v;
        }
}

void dump(const std::string &cs, const DatSpec &s, int nd, int g) {
  int e[3];
  ext_of(s, nd, e);
  const int ts = s.type == "float" ? 4 : 8;
  std::vector<char> out;
  for (int z = 0; z < e[2]; z++)
    for (int y = 0; y < e[1]; y++)
      for (int x = 0; x < e[0]; x++)
        for (int c = 0; c < s.dim; c++) {
          const char *p = raw_at(s.dat, x, y, z, c);
          out.insert(out.end(), p, p + ts);
        }
  drv_write("halo_" + cs + "_" + s.name + "_g" + std::to_string(g), out.data(), out.size());
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: ref_halo SPEC OUTDIR\n");
    return 2;
  }
  g_outdir = argv[2];
  ops_init(1, argv, 1);
  std::ifstream in(argv[1]);
  std::string line, cs;
  int nd = 0, ncase = 0, g = 0;
  std::vector<DatSpec> dats;
  std::map<std::string, int> byname;
  std::vector<ops_halo> halos;
  auto flush_group = [&]() {
    if (halos.empty()) return;
    ops_halo_group grp = ops_decl_halo_group((int)halos.size(), halos.data());
    ops_halo_transfer(grp);
    halos.clear();
    g++;
    for (auto &s : dats) dump(cs, s, nd, g);
  };
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key)) continue;
    if (key == "case") {
      ls >> cs >> nd;
      dats.clear();
      byname.clear();
      g = 0;
      ncase++;
    } else if (key == "dat") {
      DatSpec s;
      ls >> s.name >> s.type >> s.dim;
      for (int &v : s.size) ls >> v;
      for (int &v : s.dm) ls >> v;
      for (int &v : s.dp) ls >> v;
      ops_block blk = ops_decl_block(nd, (cs + "_" + s.name).c_str());
      int base[3] = {0, 0, 0};
      if (s.type == "float") {
        float *none = nullptr;
        s.dat = ops_decl_dat(blk, s.dim, s.size, base, s.dm, s.dp, none, "float", s.name.c_str());
      } else {
        double *none = nullptr;
        s.dat = ops_decl_dat(blk, s.dim, s.size, base, s.dm, s.dp, none, "double", s.name.c_str());
      }
      fill(s, nd, 100ULL * ncase + dats.size());
      byname[s.name] = (int)dats.size();
      dats.push_back(s);
    } else if (key == "group") {
      flush_group();
      if (g == 0)
        for (auto &s : dats) dump(cs, s, nd, 0);
    } else if (key == "halo") {
      std::string f, t;
      ls >> f >> t;
      int v[15];
      for (int &x : v) ls >> x;
      halos.push_back(ops_decl_halo(dats[byname.at(f)].dat, dats[byname.at(t)].dat, v, v + 3, v + 6, v + 9, v + 12));
    } else if (key == "end") {
      flush_group();
    }
  }
  ops_exit();
  return 0;
}
