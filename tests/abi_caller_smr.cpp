// abi_caller_smr.cpp -- a C++ program that drives a statically refined mesh through
// include/athena_gpu.h the way an Athena++ build with <mesh> refinement = static would:
// the MeshBlock list of the tree (level and logical location per block, in gid order, as
// Mesh::loclist holds it), host registers in AthenaArray layout copied in per block,
// Mesh::Initialize, main-loop cycles, registers copied out.  Built with g++ by
// tests/test_gpu_abi_caller.py, which writes the input file and checks the output against
// the oracle's restatement of the multilevel time stepping.
//
// usage: abi_caller_smr <input> <output>
// input (native endianness): int32 nx[3] (root grid), bnx[3], ng, xorder, rsolver,
//   ncycles, nblocks, root_level; float64 xmin[3], xmax[3], gamma, cfl; int64
//   locs[4 * nblocks] = (level, lx1, lx2, lx3) per block; then per block the registers
//   u, b1, b2, b3.
// output: per block u, w, bcc, b1, b2, b3 after ncycles VL2 cycles; float64 time, dt.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "athena_gpu.h"

namespace {

void Check(int rc, const char *what) {
  if (rc != 0) {   // the reference's ATHENA_ERROR: fatal, never TaskStatus::fail
    std::fprintf(stderr, "### FATAL ERROR in %s: %s\n", what, ath_gpu_last_error());
    std::exit(3);
  }
}

struct Reader {
  FILE *f;
  template <class T> void get(T *p, size_t n) {
    if (std::fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short input file");
  }
};

struct Writer {
  FILE *f;
  template <class T> void put(const T *p, size_t n) {
    if (std::fwrite(p, sizeof(T), n, f) != n) throw std::runtime_error("write failed");
  }
};

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
    return 2;
  }
  FILE *fi = std::fopen(argv[1], "rb");
  FILE *fo = std::fopen(argv[2], "wb");
  if (!fi || !fo) return 2;
  Reader in{fi};
  Writer out{fo};

  ath_mesh_desc d{};
  int ncycles = 0, nblocks = 0, root_level = 0;
  in.get(d.nx, 3);
  in.get(d.bnx, 3);
  in.get(&d.ng, 1);
  in.get(&d.xorder, 1);
  in.get(&d.rsolver, 1);
  in.get(&ncycles, 1);
  in.get(&nblocks, 1);
  in.get(&root_level, 1);
  in.get(d.xmin, 3);
  in.get(d.xmax, 3);
  in.get(&d.gamma, 1);
  in.get(&d.cfl, 1);
  d.dfloor = d.pfloor = -1.0;                      // reference defaults
  d.sigma_max = d.beta_min = d.gamma_max = -1.0;
  std::vector<long long> locs(4 * static_cast<size_t>(nblocks));   // Mesh::loclist
  in.get(locs.data(), locs.size());

  // GpuCreateRefined: the device mesh over the tree's block list, then the registers
  ath_mesh_t mesh = nullptr;
  Check(ath_gpu_smr_mesh_create(&d, locs.data(), nblocks, root_level, &mesh),
        "ath_gpu_smr_mesh_create");
  int nb = 0;
  Check(ath_gpu_nblocks(mesh, &nb), "ath_gpu_nblocks");
  if (nb != nblocks) Check(1, "block count");
  long ncell = 0, nf1 = 0, nf2 = 0, nf3 = 0, nbcc = 0;
  Check(ath_gpu_block_size(mesh, ATH_U, &ncell), "size u");
  Check(ath_gpu_block_size(mesh, ATH_B1, &nf1), "size b1");
  Check(ath_gpu_block_size(mesh, ATH_B2, &nf2), "size b2");
  Check(ath_gpu_block_size(mesh, ATH_B3, &nf3), "size b3");
  Check(ath_gpu_block_size(mesh, ATH_BCC, &nbcc), "size bcc");
  for (int gid = 0; gid < nb; ++gid) {               // gid is the list position
    std::vector<double> u(ncell), b1(nf1), b2(nf2), b3(nf3);
    for (std::vector<double> *a : {&u, &b1, &b2, &b3}) in.get(a->data(), a->size());
    Check(ath_gpu_set_block(mesh, gid, ATH_U, u.data()), "set u");
    Check(ath_gpu_set_block(mesh, gid, ATH_B1, b1.data()), "set b1");
    Check(ath_gpu_set_block(mesh, gid, ATH_B2, b2.data()), "set b2");
    Check(ath_gpu_set_block(mesh, gid, ATH_B3, b3.data()), "set b3");
  }
  Check(ath_gpu_initialize(mesh, 1.0e300), "ath_gpu_initialize");   // exchange, W(U), dt

  // ---- main loop -------------------------------------------------------------------
  int done = 0;
  Check(ath_gpu_cycle(mesh, ncycles, 1.0e300, &done), "ath_gpu_cycle");
  if (done != ncycles) Check(1, "cycle count");
  double time = 0.0, dt = 0.0;
  int ncycle = 0;
  Check(ath_gpu_get_clock(mesh, &time, &dt, &ncycle), "ath_gpu_get_clock");

  for (int gid = 0; gid < nb; ++gid) {
    std::vector<double> u(ncell), w(ncell), bcc(nbcc), b1(nf1), b2(nf2), b3(nf3);
    Check(ath_gpu_get_block(mesh, gid, ATH_U, u.data()), "get u");
    Check(ath_gpu_get_block(mesh, gid, ATH_W, w.data()), "get w");
    Check(ath_gpu_get_block(mesh, gid, ATH_BCC, bcc.data()), "get bcc");
    Check(ath_gpu_get_block(mesh, gid, ATH_B1, b1.data()), "get b1");
    Check(ath_gpu_get_block(mesh, gid, ATH_B2, b2.data()), "get b2");
    Check(ath_gpu_get_block(mesh, gid, ATH_B3, b3.data()), "get b3");
    for (std::vector<double> *a : {&u, &w, &bcc, &b1, &b2, &b3}) out.put(a->data(), a->size());
  }
  out.put(&time, 1);
  out.put(&dt, 1);
  Check(ath_gpu_mesh_destroy(mesh), "ath_gpu_mesh_destroy");
  std::fclose(fi);
  std::fclose(fo);
  std::printf("abi_caller_smr: %d MeshBlocks, %d cycles, time %.17g\n", nb, ncycles, time);
  return 0;
}
