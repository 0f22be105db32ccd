// raytracer_overlap_estimated (query/src/applications/raytracer_overlap_estimated_main.cpp).
#include "overlap_hash_app.hpp"
int main(int argc, char** argv) { return rs3d_cli::runOverlapHash(argc, argv, false); }
