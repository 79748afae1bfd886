// host/run_main.cpp -- `run`: subgraph enumeration of a preset pattern
// (app_gmatch/main.cu:13-20)
//   run -dg G.bin -q 0     (triangles)
//   run -dg G.bin -q 24    (4-cliques)
#include "worker.hpp"

int main(int argc, char** argv) { return g2w::run_worker(argc, argv, g2w::App::GM); }
