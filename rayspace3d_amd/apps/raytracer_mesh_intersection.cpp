// raytracer_mesh_intersection: the intersection join without estimation (named in the
// reference README, README.md:72,127; semantics README.md:152-153).
#include "intersection_app.hpp"
int main(int argc, char** argv) { return rs3d_cli::runIntersection(argc, argv, false); }
