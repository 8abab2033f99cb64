// Seeded Lloyd-relaxed Voronoi meshes of a disk or an annulus (placeholder).
#include "common.h"

namespace vemd {

Mesh* mesh_generate(const vemd_mesh_opts& o) {
    (void)o;
    throw Error(VEMD_ERR_ARG, "mesh generator not built yet");
}

}  // namespace vemd
