// raytracer_intersection_estimated
// (query/src/applications/raytracer_intersection_estimated_main.cpp).
#include "intersection_app.hpp"
int main(int argc, char** argv) { return rs3d_cli::runIntersection(argc, argv, true); }
