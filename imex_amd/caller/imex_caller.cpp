// imex_caller — a reference-side C++ program that drives libimex_amd.so through include/imex.h the way a
// maintainer of the reference's main.cpp would: the host keeps its Grid / Array buffers in the
// reference layout and hands them to the library.
//
//   imex_caller step <dir> <nsteps>
//       The whole-step drop-in (replaces the body of the do { ... } while loop, main.cpp:334-561):
//       reads <dir>/params.bin (an imex_params), u_pri.bin, bx.bin, by.bin, runs imex_step nsteps
//       times one step at a time (the reference's loop, main.cpp:334), then writes out_u_con.bin,
//       out_u_pri.bin, out_bx.bin, out_by.bin, out_scalars.bin (imex_scalars), out_stats.bin.
//
//   imex_caller stage <dir> <subsystem>
//       One subsystem replaced in place (e.g. PreSubsystem_Pressure at main.cpp:485,
//       declarations.hpp:74-96): reads params.bin, scalars.bin and the stage working set
//       (u_con, Qe, Qs, bx, by, bx_qs, by_qs, fx, fy .bin), runs imex_set_stage, the named entry
//       point (explicit_conv, mag_update_b, explicit_press, pre_subsystem_pressure, explicit_energy,
//       pre_subsystem_energy, ct_update), imex_get_stage, and writes the working set back as out_*.bin.
//
// Exit status: 0 on success (warnings are printed), 1 on a library error, 2 on a usage or file error.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "imex.h"

namespace {

constexpr int kGhost = 3;   // ghost layers of the reference's Grid / Array (include/imex.h layout note)
constexpr int kNvar = 8;

bool read_file(const std::string& path, void* dst, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) { std::fprintf(stderr, "imex_caller: cannot open %s\n", path.c_str()); return false; }
  const size_t got = std::fread(dst, 1, bytes, f);
  const bool extra = std::fgetc(f) != EOF;
  std::fclose(f);
  if (got != bytes || extra) {
    std::fprintf(stderr, "imex_caller: %s holds %s than the %zu bytes expected\n", path.c_str(),
                 got != bytes ? "fewer" : "more", bytes);
    return false;
  }
  return true;
}

bool write_file(const std::string& path, const void* src, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { std::fprintf(stderr, "imex_caller: cannot create %s\n", path.c_str()); return false; }
  const size_t put = std::fwrite(src, 1, bytes, f);
  std::fclose(f);
  return put == bytes;
}

// Host buffers of one problem size: Grid = (nx+6)(ny+6) cells x 8 slots, Array = (nx+6)(ny+6).
struct Sizes {
  size_t grid, array;
  explicit Sizes(const imex_params& p)
      : grid((size_t)(p.nx + 2 * kGhost) * (p.ny + 2 * kGhost) * kNvar),
        array((size_t)(p.nx + 2 * kGhost) * (p.ny + 2 * kGhost)) {}
};

int fail(const char* what) {
  std::fprintf(stderr, "imex_caller: %s failed: %s\n", what, imex_last_error());
  return 1;
}

int run_step(const std::string& dir, int nsteps) {
  imex_params prm;
  if (!read_file(dir + "/params.bin", &prm, sizeof prm)) return 2;
  const Sizes n(prm);
  std::vector<double> u_pri(n.grid), u_con(n.grid), bx(n.array), by(n.array);
  if (!read_file(dir + "/u_pri.bin", u_pri.data(), n.grid * sizeof(double)) ||
      !read_file(dir + "/bx.bin", bx.data(), n.array * sizeof(double)) ||
      !read_file(dir + "/by.bin", by.data(), n.array * sizeof(double)))
    return 2;

  // once, after u_initialization / bound_cond_staggered (main.cpp:240-242)
  imex_ctx* ctx = nullptr;
  if (imex_create(&prm, &ctx) < 0) return fail("imex_create");
  int rc = imex_set_state(ctx, u_pri.data(), bx.data(), by.data(), 0.0, 0, 0.0);
  if (rc < 0) { imex_destroy(ctx); return fail("imex_set_state"); }

  // the reference's loop, one imex_step per pass (main.cpp:334-561)
  imex_stats st;
  int warnings = 0;
  for (int s = 0; s < nsteps; s++) {
    rc = imex_step(ctx, 1, &st);
    if (rc < 0) { imex_destroy(ctx); return fail("imex_step"); }
    warnings |= rc;
  }
  imex_scalars sc;
  rc = imex_get_state(ctx, u_con.data(), u_pri.data(), bx.data(), by.data(), &sc);
  if (rc < 0) { imex_destroy(ctx); return fail("imex_get_state"); }
  imex_destroy(ctx);
  if (warnings) std::printf("warnings 0x%x\n", warnings);
  std::printf("t %.17g timestep %d div_b %.3e\n", sc.t, sc.timestep, st.div_b);
  const bool ok = write_file(dir + "/out_u_con.bin", u_con.data(), n.grid * sizeof(double)) &&
                  write_file(dir + "/out_u_pri.bin", u_pri.data(), n.grid * sizeof(double)) &&
                  write_file(dir + "/out_bx.bin", bx.data(), n.array * sizeof(double)) &&
                  write_file(dir + "/out_by.bin", by.data(), n.array * sizeof(double)) &&
                  write_file(dir + "/out_scalars.bin", &sc, sizeof sc) &&
                  write_file(dir + "/out_stats.bin", &st, sizeof st);
  return ok ? 0 : 2;
}

int run_stage(const std::string& dir, const std::string& name) {
  typedef int (*Entry)(imex_ctx*);
  struct Named { const char* name; Entry fn; };
  static const Named kEntries[] = {
      {"explicit_conv", imex_explicit_conv},           {"mag_update_b", imex_mag_update_b},
      {"explicit_press", imex_explicit_press},         {"pre_subsystem_pressure", imex_pre_subsystem_pressure},
      {"explicit_energy", imex_explicit_energy},       {"pre_subsystem_energy", imex_pre_subsystem_energy},
      {"ct_update", imex_ct_update}};
  Entry fn = nullptr;
  for (const Named& e : kEntries)
    if (name == e.name) fn = e.fn;
  if (!fn) { std::fprintf(stderr, "imex_caller: unknown subsystem %s\n", name.c_str()); return 2; }

  imex_params prm;
  imex_scalars sc;
  if (!read_file(dir + "/params.bin", &prm, sizeof prm) || !read_file(dir + "/scalars.bin", &sc, sizeof sc)) return 2;
  const Sizes n(prm);
  static const char* kGrids[] = {"u_con", "Qe", "Qs"};
  static const char* kArrays[] = {"bx", "by", "bx_qs", "by_qs", "fx", "fy"};
  std::vector<double> g[3], a[6];
  for (int k = 0; k < 3; k++) {
    g[k].resize(n.grid);
    if (!read_file(dir + "/" + kGrids[k] + ".bin", g[k].data(), n.grid * sizeof(double))) return 2;
  }
  for (int k = 0; k < 6; k++) {
    a[k].resize(n.array);
    if (!read_file(dir + "/" + kArrays[k] + ".bin", a[k].data(), n.array * sizeof(double))) return 2;
  }
  // the reference's own buffers, handed over as they are (INTEGRATION: replace one subsystem)
  imex_stage_host h = {g[0].data(), g[1].data(), g[2].data(), a[0].data(), a[1].data(),
                       a[2].data(), a[3].data(), a[4].data(), a[5].data()};
  imex_ctx* ctx = nullptr;
  if (imex_create(&prm, &ctx) < 0) return fail("imex_create");
  int rc = imex_set_stage(ctx, &h, &sc);
  if (rc >= 0) rc = fn(ctx);
  if (rc >= 0) rc = imex_get_stage(ctx, &h, &sc);
  if (rc < 0) { std::fprintf(stderr, "imex_caller: %s: %s\n", name.c_str(), imex_last_error()); imex_destroy(ctx); return 1; }
  imex_destroy(ctx);
  if (rc > 0) std::printf("warnings 0x%x\n", rc);
  bool ok = write_file(dir + "/out_scalars.bin", &sc, sizeof sc);
  for (int k = 0; k < 3 && ok; k++) ok = write_file(dir + "/out_" + kGrids[k] + ".bin", g[k].data(), n.grid * sizeof(double));
  for (int k = 0; k < 6 && ok; k++) ok = write_file(dir + "/out_" + kArrays[k] + ".bin", a[k].data(), n.array * sizeof(double));
  return ok ? 0 : 2;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 4 && std::strcmp(argv[1], "step") == 0) {
    const int nsteps = std::atoi(argv[3]);
    if (nsteps < 0) { std::fprintf(stderr, "imex_caller: bad step count\n"); return 2; }
    return run_step(argv[2], nsteps);
  }
  if (argc == 4 && std::strcmp(argv[1], "stage") == 0) return run_stage(argv[2], argv[3]);
  std::fprintf(stderr, "usage: imex_caller step <dir> <nsteps> | imex_caller stage <dir> <subsystem>\n");
  return 2;
}
