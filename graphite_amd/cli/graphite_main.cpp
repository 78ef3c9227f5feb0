// graphite <case.json>: runs a steady Euler case file (the reference's apps/main.cu CLI) on the
// device solver through libgraphite_hip.so.
#include <cstdio>

#include "graphite_hip.h"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::printf("Error, no file provided.\nusage: graphite <case.json>\n");
        return 1;
    }
    gr_solve_result r{};
    const int rc = gr_case_run(argv[1], 1, &r);
    if (rc != GR_OK) {
        std::fprintf(stderr, "graphite: %s\n", gr_last_error());
        return rc;
    }
    return r.converged ? 0 : 7;
}
