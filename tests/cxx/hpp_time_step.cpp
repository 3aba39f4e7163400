// Drives the C++ host mirror (include/exwave_amd.hpp) the way the reference's
// driver drives its classes (source/explicit_wave.cc:549-799): mesh, materials,
// WaveEquationOperation<3,4>::setup, project_initial_field, then one
// perform_time_step of LowStorageRK45Reg2 (argv[1] = "rk") or of
// ArbitraryHighOrderDG over WaveEquationOperationADER<3,4> ("ader"), and writes
// the initial and final state as raw doubles to argv[2].
//
// Mesh: 3^3 cells of the reference's sine-deformed unit cube (factor 0.1),
// boundary id 2, material_id = cell % 2 with materials {rho 1, c 1} and
// {rho 2, c 4}; initial data: modes (1,1,1) + (2,3,1).
#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

#include "../../include/exwave_amd.hpp"

using namespace exwave_amd;

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s rk|ader out.bin\n", argv[0]);
    return 2;
  }
  try {
    Context ctx(0);
    Mesh mesh = Mesh::hypercube(3, 3, 0, EXW_TRANSFORM_REFERENCE_SINE, 0.1, 2);
    for (int c = 0; c < mesh.n_cells; ++c) mesh.material_id[c] = c % 2;
    std::vector<Material> mats(2);
    mats[1].density = 2.0;
    mats[1].speed = 4.0;
    exw_modes modes{};
    modes.n_modes = 2;
    modes.amp[0] = modes.amp[1] = 1.0;
    const int m[2][3] = {{1, 1, 1}, {2, 3, 1}};
    std::memcpy(modes.m, m, sizeof(m));
    modes.rho = modes.c = 1.0;

    TimeControl time_control;
    const bool ader = std::strcmp(argv[1], "ader") == 0;
    std::unique_ptr<WaveEquationOperation<3, 4>> op;
    if (ader)
      op.reset(new WaveEquationOperationADER<3, 4>(ctx, time_control, false));
    else
      op.reset(new WaveEquationOperation<3, 4>(ctx, time_control));
    op->setup(mesh, mats);
    // compute_time_step_size (source/explicit_wave.cc:529-545), cfl 0.2 / k^1.5
    const double dt = 0.2 / std::pow(4.0, 1.5) * op->min_vertex_distance_over_c();
    time_control.setup(1.0, 0.1, dt);
    time_control.set_time_step(dt);  // the unrounded step (source/explicit_wave.cc:222)

    const int64_t N = op->n_dofs();
    DeviceVector solution(ctx, N), tmp(ctx, N);
    op->project_initial_field(solution, modes);
    const std::vector<double> u0 = solution.download();
    std::unique_ptr<ExplicitIntegrator<DeviceVector, WaveEquationOperationBase<3>>> integ;
    if (ader)
      integ.reset(new ArbitraryHighOrderDG<DeviceVector, WaveEquationOperationBase<3>>(false));
    else
      integ.reset(new LowStorageRK45Reg2<DeviceVector, WaveEquationOperationBase<3>>());
    // the time loop's swap and step (source/explicit_wave.cc:784-786)
    solution.swap(tmp);
    integ->perform_time_step(tmp, solution, time_control.get_time_step(), *op);
    time_control.advance_time_step();
    const std::vector<double> u1 = solution.download();

    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 3;
    const double hdr[2] = {(double)N, dt};
    std::fwrite(hdr, sizeof(double), 2, f);
    std::fwrite(u0.data(), sizeof(double), u0.size(), f);
    std::fwrite(u1.data(), sizeof(double), u1.size(), f);
    std::fclose(f);
    std::printf("%s step done: %lld DoFs, dt = %.10g\n", argv[1], (long long)N, dt);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
