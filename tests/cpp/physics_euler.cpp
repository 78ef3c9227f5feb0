// A host program written against the reference's call surface (include/graphite/graphite.hpp
// mirrors include/mesh.cuh, include/meshio.cuh, include/fvmfields.cuh, include/fvmops.cuh and
// include/models/euler.cuh), following PhysicsEuler's set-up and step protocol
// (src/models/euler.cu:20-56 and :99-175): convert the su2 mesh, read the .fvm file, create
// the fields, initialise to freestream, then per iteration apply the boundary conditions in
// order, FvmEulerStepPrepare, FvmEulerStep("steady") and FvmFieldReduceSum of the error field.
//
// usage: physics_euler <mesh.su2> <mesh.fvm> <iterations> <cfl> <mach_x> <mach_y> <mach_z>
//                      <history.txt> <state.bin> <region>:<slip|farfield> ...
// Writes one line per iteration (the residual sum, %.17g) and the final state (rho, rhou x3,
// rhoe per cell and ghost, binary64).  Exit status 0 on success, 2 on a library error (its
// message on stderr), 1 on bad arguments.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

#include "graphite/graphite.hpp"

int main(int argc, char **argv) {
    if (argc < 11) {
        std::fprintf(stderr, "usage: %s su2 fvm iterations cfl mx my mz history state region:kind...\n", argv[0]);
        return 1;
    }
    const int n_iter = std::atoi(argv[3]);
    const scalar cfl = std::atof(argv[4]);
    const vector mach = make_vector(std::atof(argv[5]), std::atof(argv[6]), std::atof(argv[7]));
    const scalar pressure = 1.0, temperature = 1.0 / 1.4;
    std::vector<std::pair<std::string, std::string>> bcs;
    for (int a = 10; a < argc; ++a) {
        std::string s = argv[a];
        const size_t c = s.find(':');
        if (c == std::string::npos) return 1;
        bcs.emplace_back(s.substr(0, c), s.substr(c + 1));
    }
    try {
        if (cuMeshFromSu2(argv[1], argv[2], 8)) {
            std::fprintf(stderr, "Error during mesh conversion\n");
            return 2;
        }
        cuMesh *mesh = cuMeshFromFile(argv[2], 8);
        const size_t var_size = cuMeshGetVarSize(mesh);
        FvmField *rho = FvmFieldNew(sizeof(scalar), var_size, "rho", 1);
        FvmField *rhou = FvmFieldNew(sizeof(vector), var_size, "rhou", 1);
        FvmField *rhoe = FvmFieldNew(sizeof(scalar), var_size, "rhoe", 1);
        FvmField *error = FvmFieldNew(sizeof(scalar), var_size, "error", 1);
        FvmEulerInit(mesh, rho, rhou, rhoe, mach, pressure, temperature);

        FILE *hist = std::fopen(argv[8], "w");
        if (!hist) return 1;
        scalar err0 = 0;
        for (int iter = 0; iter < n_iter; ++iter) {
            for (const auto &bc : bcs) {
                if (bc.second == "slip") FvmEulerSlipBc(mesh, rho, rhou, rhoe, bc.first.c_str());
                else if (bc.second == "farfield")
                    FvmEulerInletOutletBc(mesh, rho, rhou, rhoe, mach, pressure, temperature, bc.first.c_str());
            }
            FvmEulerStepPrepare(mesh, rho, rhou, rhoe);
            FvmEulerStep(mesh, rho, rhou, rhoe, cfl, "steady", NULL, error);
            const scalar err = FvmFieldReduceSum(mesh, error);
            if (iter == 0) err0 = err;
            std::fprintf(hist, "%.17g %.17g\n", err, err / err0);
        }
        std::fclose(hist);

        VarGpuToCpu(rho);
        VarGpuToCpu(rhou);
        VarGpuToCpu(rhoe);
        FILE *st = std::fopen(argv[9], "wb");
        if (!st) return 1;
        for (size_t i = 0; i < var_size; ++i) {
            const double *r = (const double *)FvmFieldValueIndex(rho, i);
            const double *u = (const double *)FvmFieldValueIndex(rhou, i);
            const double *e = (const double *)FvmFieldValueIndex(rhoe, i);
            const double rec[5] = {r[0], u[0], u[1], u[2], e[0]};
            std::fwrite(rec, sizeof(double), 5, st);
        }
        std::fclose(st);

        // a field must stay with the mesh it was first used with
        cuMesh *other = cuMeshFromFile(argv[2], 8);
        bool refused = false;
        try {
            FvmEulerSlipBc(other, rho, rhou, rhoe, bcs.front().first.c_str());
        } catch (const graphite_error &e) {
            refused = e.code == GR_ERR_INVALID;
        }
        // with two meshes of the same size, an unbound field cannot pick one
        FvmField *loose = FvmFieldNew(sizeof(scalar), var_size, "loose", 0);
        bool ambiguous = false;
        try {
            VarGpuToCpu(loose);
        } catch (const graphite_error &e) {
            ambiguous = e.code == GR_ERR_INVALID;
        }
        std::printf("binding_checks %d %d\n", refused ? 1 : 0, ambiguous ? 1 : 0);
        for (FvmField *f : {rho, rhou, rhoe, error, loose}) FvmFieldFree(f);
        cuMeshFree(other);
        cuMeshFree(mesh);
    } catch (const graphite_error &e) {
        std::fprintf(stderr, "graphite error %d: %s\n", e.code, e.what());
        return 2;
    }
    return 0;
}
