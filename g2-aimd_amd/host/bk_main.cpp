// host/bk_main.cpp -- `bk`: maximal clique enumeration (app_BK/main.cu:13-21)
//   bk -dg G.bin -h 2 -pn 4 -cn 4 -t 4 -qs 4 -ss 1
#include "worker.hpp"

int main(int argc, char** argv) { return g2w::run_worker(argc, argv, g2w::App::BK); }
