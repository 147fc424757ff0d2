// abi_caller.cpp -- a C++ program that drives libathena_gpu.so through include/athena_gpu.h
// the way an Athena++ build would bind it: host registers in AthenaArray layout (one
// contiguous array per register, i fastest), the mesh-level calls of the task list, then
// the per-MeshBlock functions over device arrays that the caller owns.  Built with g++
// by tests/test_gpu_abi_caller.py, which writes the input file and checks the output
// against the oracle.
//
// usage: abi_caller <input> <output>
// input (native endianness): int32 nx[3], bnx[3], ng, xorder, rsolver, relativistic,
//   bc[6], ncycles; float64 xmin[3], xmax[3], gamma, cfl; then per MeshBlock the
//   registers u, b1, b2, b3, w, w1; then block 0's coordinates x1f, x2f, x3f, x1v, x2v,
//   x3v, dx1f, dx2f, dx3f.
// output: per MeshBlock u, w, bcc, b1, b2, b3 after ncycles VL2 cycles; float64 time, dt;
//   then, starting from block 0's final registers: u, b1, b2, b3 after
//   CalculateFluxes(xorder) + ComputeCornerE + CT(dt/2) + AddFluxDivergence(dt/2), w and
//   bcc after ConservedToPrimitive over the whole block, and its NewBlockTimeStep.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "athena_gpu.h"

namespace {

void Check(int rc, const char *what) {
  if (rc != 0) {   // the reference's ATHENA_ERROR: fatal, never TaskStatus::fail
    std::fprintf(stderr, "### FATAL ERROR in %s: %s\n", what, ath_gpu_last_error());
    std::exit(3);
  }
}

void HipCheck(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "### HIP error in %s: %s\n", what, hipGetErrorString(e));
    std::exit(4);
  }
}

// An AthenaArray<Real> stand-in: contiguous storage, (n, k, j, i) with i fastest.
struct HostArray {
  std::vector<double> v;
  explicit HostArray(long n = 0) : v(static_cast<size_t>(n), 0.0) {}
  double *data() { return v.data(); }
};

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

double *DeviceCopy(const HostArray &h) {
  double *d = nullptr;
  HipCheck(hipMalloc(reinterpret_cast<void **>(&d), sizeof(double) * h.v.size()), "hipMalloc");
  HipCheck(hipMemcpy(d, h.v.data(), sizeof(double) * h.v.size(), hipMemcpyHostToDevice),
           "hipMemcpy");
  return d;
}

double *DeviceZeros(long n) {
  double *d = nullptr;
  HipCheck(hipMalloc(reinterpret_cast<void **>(&d), sizeof(double) * n), "hipMalloc");
  HipCheck(hipMemset(d, 0, sizeof(double) * n), "hipMemset");
  return d;
}

void ToHost(HostArray &h, const double *d) {
  HipCheck(hipMemcpy(h.v.data(), d, sizeof(double) * h.v.size(), hipMemcpyDeviceToHost),
           "hipMemcpy");
}

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

  // ---- mesh set-up (Mesh::Initialize after the problem generator) ------------------
  ath_mesh_desc d{};
  int ncycles = 0;
  in.get(d.nx, 3);
  in.get(d.bnx, 3);
  in.get(&d.ng, 1);
  in.get(&d.xorder, 1);
  in.get(&d.rsolver, 1);
  in.get(&d.relativistic, 1);
  in.get(d.bc, 6);
  in.get(&ncycles, 1);
  in.get(d.xmin, 3);
  in.get(d.xmax, 3);
  in.get(&d.gamma, 1);
  in.get(&d.cfl, 1);
  d.dfloor = d.pfloor = -1.0;                      // reference defaults
  d.sigma_max = d.beta_min = d.gamma_max = -1.0;
  ath_mesh_t mesh = nullptr;
  Check(ath_gpu_mesh_create(&d, &mesh), "ath_gpu_mesh_create");
  int nb = 0;
  Check(ath_gpu_nblocks(mesh, &nb), "ath_gpu_nblocks");
  const int c1 = d.bnx[0] + 2 * d.ng, c2 = d.bnx[1] + 2 * d.ng, c3 = d.bnx[2] + 2 * d.ng;
  const long ncell = static_cast<long>(c1) * c2 * c3;
  const long nf1 = static_cast<long>(c3) * c2 * (c1 + 1);
  const long nf2 = static_cast<long>(c3) * (c2 + 1) * c1;
  const long nf3 = static_cast<long>(c3 + 1) * c2 * c1;
  const long ne1 = static_cast<long>(c3 + 1) * (c2 + 1) * c1;
  const long ne2 = static_cast<long>(c3 + 1) * c2 * (c1 + 1);
  const long ne3 = static_cast<long>(c3) * (c2 + 1) * (c1 + 1);
  for (int gid = 0; gid < nb; ++gid) {
    HostArray u(5 * ncell), b1(nf1), b2(nf2), b3(nf3), w(5 * ncell), w1(5 * ncell);
    for (HostArray *a : {&u, &b1, &b2, &b3, &w, &w1}) in.get(a->data(), a->v.size());
    Check(ath_gpu_set_block(mesh, gid, ATH_U, u.data()), "set u");
    Check(ath_gpu_set_block(mesh, gid, ATH_B1, b1.data()), "set b1");
    Check(ath_gpu_set_block(mesh, gid, ATH_B2, b2.data()), "set b2");
    Check(ath_gpu_set_block(mesh, gid, ATH_B3, b3.data()), "set b3");
    Check(ath_gpu_set_block(mesh, gid, ATH_W, w.data()), "set w");
    Check(ath_gpu_set_block(mesh, gid, ATH_W1, w1.data()), "set w1");
  }
  std::vector<std::vector<double>> coords;
  const int ncs[3] = {c1, c2, c3};
  for (int q = 0; q < 3; ++q) coords.emplace_back(ncs[q] + 1);
  for (int q = 0; q < 3; ++q) coords.emplace_back(ncs[q]);
  for (int q = 0; q < 3; ++q) coords.emplace_back(ncs[q]);
  for (auto &c : coords) in.get(c.data(), c.size());
  Check(ath_gpu_initialize(mesh, 1.0e300), "ath_gpu_initialize");

  // ---- task list: one batched call per stage ---------------------------------------
  for (int n = 0; n < ncycles; ++n) {
    Check(ath_gpu_vl2_stage(mesh, 1), "stage 1");
    Check(ath_gpu_vl2_stage(mesh, 2), "stage 2");   // also ncycle++, time += dt, new dt
  }
  double time = 0.0, dt = 0.0;
  int ncycle = 0;
  Check(ath_gpu_get_clock(mesh, &time, &dt, &ncycle), "ath_gpu_get_clock");
  HostArray u0(5 * ncell), w0(5 * ncell), bcc0(3 * ncell), b10(nf1), b20(nf2), b30(nf3);
  for (int gid = 0; gid < nb; ++gid) {
    HostArray u(5 * ncell), w(5 * ncell), bcc(3 * ncell), b1(nf1), b2(nf2), b3(nf3);
    Check(ath_gpu_get_block(mesh, gid, ATH_U, u.data()), "get u");
    Check(ath_gpu_get_block(mesh, gid, ATH_W, w.data()), "get w");
    Check(ath_gpu_get_block(mesh, gid, ATH_BCC, bcc.data()), "get bcc");
    Check(ath_gpu_get_block(mesh, gid, ATH_B1, b1.data()), "get b1");
    Check(ath_gpu_get_block(mesh, gid, ATH_B2, b2.data()), "get b2");
    Check(ath_gpu_get_block(mesh, gid, ATH_B3, b3.data()), "get b3");
    for (HostArray *a : {&u, &w, &bcc, &b1, &b2, &b3}) out.put(a->data(), a->v.size());
    if (gid == 0) { u0 = u; w0 = w; bcc0 = bcc; b10 = b1; b20 = b2; b30 = b3; }
  }
  out.put(&time, 1);
  out.put(&dt, 1);
  Check(ath_gpu_mesh_destroy(mesh), "ath_gpu_mesh_destroy");

  // ---- per-MeshBlock functions over the caller's device arrays ---------------------
  ath_block_desc bd{};
  for (int q = 0; q < 3; ++q) bd.nx[q] = d.bnx[q];
  bd.ng = d.ng;
  bd.x1f = coords[0].data(); bd.x2f = coords[1].data(); bd.x3f = coords[2].data();
  bd.x1v = coords[3].data(); bd.x2v = coords[4].data(); bd.x3v = coords[5].data();
  bd.dx1f = coords[6].data(); bd.dx2f = coords[7].data(); bd.dx3f = coords[8].data();
  bd.rsolver = d.rsolver;
  bd.relativistic = d.relativistic;
  bd.gamma = d.gamma;
  bd.dfloor = bd.pfloor = -1.0;
  bd.sigma_max = bd.beta_min = bd.gamma_max = -1.0;
  bd.cfl = d.cfl;
  ath_block_t blk = nullptr;
  Check(ath_gpu_block_create(&bd, &blk), "ath_gpu_block_create");
  double *du = DeviceCopy(u0), *dw = DeviceCopy(w0), *dbcc = DeviceCopy(bcc0);
  double *db1 = DeviceCopy(b10), *db2 = DeviceCopy(b20), *db3 = DeviceCopy(b30);
  ath_block_regs r{};
  r.flux[0] = DeviceZeros(5 * nf1); r.flux[1] = DeviceZeros(5 * nf2);
  r.flux[2] = DeviceZeros(5 * nf3);
  r.e3_x1f = DeviceZeros(nf1); r.e2_x1f = DeviceZeros(nf1);
  r.e1_x2f = DeviceZeros(nf2); r.e3_x2f = DeviceZeros(nf2);
  r.e2_x3f = DeviceZeros(nf3); r.e1_x3f = DeviceZeros(nf3);
  r.wght[0] = DeviceZeros(nf1); r.wght[1] = DeviceZeros(nf2); r.wght[2] = DeviceZeros(nf3);
  r.e[0] = DeviceZeros(ne1); r.e[1] = DeviceZeros(ne2); r.e[2] = DeviceZeros(ne3);
  double *ddt = DeviceZeros(1);
  void *stream = nullptr;
  Check(ath_gpu_block_calculate_fluxes(blk, dw, db1, db2, db3, dbcc, d.xorder, dt, &r, stream),
        "CalculateFluxes");
  Check(ath_gpu_block_compute_corner_e(blk, dw, dbcc, &r, stream), "ComputeCornerE");
  Check(ath_gpu_block_ct(blk, 0.5 * dt, &r, db1, db2, db3, stream), "CT");
  Check(ath_gpu_block_add_flux_divergence(blk, 0.5 * dt, &r, du, stream), "AddFluxDivergence");
  Check(ath_gpu_block_cons2prim(blk, du, dw, db1, db2, db3, dw, dbcc, 0, c1 - 1, 0, c2 - 1, 0,
                                c3 - 1, stream),
        "ConservedToPrimitive");
  Check(ath_gpu_block_new_dt(blk, dw, dbcc, db1, db2, db3, ddt, stream), "NewBlockTimeStep");
  HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
  ToHost(u0, du); ToHost(b10, db1); ToHost(b20, db2); ToHost(b30, db3);
  ToHost(w0, dw); ToHost(bcc0, dbcc);
  double new_dt = 0.0;
  HipCheck(hipMemcpy(&new_dt, ddt, sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
  for (HostArray *a : {&u0, &b10, &b20, &b30, &w0, &bcc0}) out.put(a->data(), a->v.size());
  out.put(&new_dt, 1);
  for (double *p : {du, dw, dbcc, db1, db2, db3, ddt, r.flux[0], r.flux[1], r.flux[2],
                    r.e3_x1f, r.e2_x1f, r.e1_x2f, r.e3_x2f, r.e2_x3f, r.e1_x3f, r.wght[0],
                    r.wght[1], r.wght[2], r.e[0], r.e[1], r.e[2]})
    HipCheck(hipFree(p), "hipFree");
  Check(ath_gpu_block_destroy(blk), "ath_gpu_block_destroy");
  std::fclose(fi);
  std::fclose(fo);
  std::printf("abi_caller: %d MeshBlocks, %d cycles, time %.17g, dt %.17g, block dt %.17g\n",
              nb, ncycle, time, dt, new_dt);
  return 0;
}
