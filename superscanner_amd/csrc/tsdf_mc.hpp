// tsdf_mc.hpp -- marching cubes over the bricks, the mesh the reference
// builds in marchingCubesHost (GPUVoxelGrid.cpp:200-251): the bricks are
// copied into an Open3D 0.12 ScalableTSDFVolume (8^3 voxels per unit, RGB8
// colour) and meshed by its ExtractTriangleMesh (:248), then
// ComputeVertexNormals (:249).  Included once, at the end of tsdf_hip.hip.
//
// The copy at GPUVoxelGrid.cpp:226-241 walks (x, y, z) with z fastest and
// stores voxel x + 8y + 64z of the brick (storage order) at Open3D index
// x + 8y + 64z, which Open3D's UniformTSDFVolume::IndexOf reads as voxel
// (z, y, x) (its index is 64 X + 8 Y + Z); the voxel it reads there is
// storage voxel (z, y, x), so the two transpositions cancel and the volume
// holds each voxel at its own (x, y, z).  Semantics of ExtractTriangleMesh
// (Open3D's own implementation of Bourke's "Polygonising a scalar field"
// with the Lorensen-Cline case tables), restated:
//   * cell (x, y, z) of unit (bx, by, bz) has corners k = 0..7 at voxel
//     (x, y, z) + offset(k) (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1)
//     (1,1,1) (0,1,1), in the same unit or the next one along an axis (a
//     missing unit reads w = 0);
//   * a corner with w == 0 drops the cell (so w = -1, a deintegration
//     remainder, counts as observed); otherwise bit k of the case is
//     sdf_k < 0; cases 0 and 255 emit nothing;
//   * a vertex per lattice edge: edge (global voxel G, axis a) from G to
//     G + e_a, position (vl/2 + vl G) in double with
//     f0 vl / (f0 + f1) added along a, f0 = |sdf(G)|, f1 = |sdf(G + e_a)|
//     (doubles); colour (f1 c0 + f0 c1) / (f0 + f1) with c = rgb / 255;
//     every cell that shares the edge shares the vertex;
//   * triangle t of the case row (e[3t], e[3t+1], e[3t+2]) is emitted as
//     (e[3t], e[3t+2], e[3t+1]), which faces the positive (free-space) side;
//   * vertex normal = sum of the unnormalised face normals (v1 - v0) x
//     (v2 - v0) of its triangles, then normalised (a zero sum stays zero).
// Open3D's vertex and triangle order follows its hash-map iteration order;
// here triangles come in brick order, cells in x + 8y + 64z order within a
// brick, and vertex ids in the order a parallel pass claims them, so the
// mesh equals the reference's as a set of vertices (each named by its edge)
// and a multiset of triangles over them.
#pragma once

namespace {

// Corner k's offset as bits x + 2y + 4z.
__constant__ uint8_t kMcCorner[8] = {0, 1, 3, 2, 4, 5, 7, 6};
// Edge e runs from corner kMcEdgeLo[e] to corner kMcEdgeHi[e] along axis
// kMcEdgeAxis[e] (the low corner first).
__constant__ uint8_t kMcEdgeLo[12] = {0, 1, 3, 0, 4, 5, 7, 4, 0, 1, 2, 3};
__constant__ uint8_t kMcEdgeHi[12] = {1, 2, 2, 3, 5, 6, 6, 7, 4, 5, 6, 7};
__constant__ uint8_t kMcEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
// Triangles of each case (bit k = corner k inside), edges in threes, -1
// terminated: the published Lorensen-Cline / Bourke table.
__constant__ int8_t kMcTri[256][16] = {
    {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 8, 3, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 1, 9, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 8, 3, 9, 8, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 2, 10, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 8, 3, 1, 2, 10, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {9, 2, 10, 0, 2, 9, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {2, 8, 3, 2, 10, 8, 10, 9, 8, -1, -1, -1, -1, -1, -1, -1},
    {3, 11, 2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 11, 2, 8, 11, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 9, 0, 2, 3, 11, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 11, 2, 1, 9, 11, 9, 8, 11, -1, -1, -1, -1, -1, -1, -1},
    {3, 10, 1, 11, 10, 3, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 10, 1, 0, 8, 10, 8, 11, 10, -1, -1, -1, -1, -1, -1, -1},
    {3, 9, 0, 3, 11, 9, 11, 10, 9, -1, -1, -1, -1, -1, -1, -1},
    {9, 8, 10, 10, 8, 11, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {4, 7, 8, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {4, 3, 0, 7, 3, 4, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 1, 9, 8, 4, 7, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {4, 1, 9, 4, 7, 1, 7, 3, 1, -1, -1, -1, -1, -1, -1, -1},
    {1, 2, 10, 8, 4, 7, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {3, 4, 7, 3, 0, 4, 1, 2, 10, -1, -1, -1, -1, -1, -1, -1},
    {9, 2, 10, 9, 0, 2, 8, 4, 7, -1, -1, -1, -1, -1, -1, -1},
    {2, 10, 9, 2, 9, 7, 2, 7, 3, 7, 9, 4, -1, -1, -1, -1},
    {8, 4, 7, 3, 11, 2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {11, 4, 7, 11, 2, 4, 2, 0, 4, -1, -1, -1, -1, -1, -1, -1},
    {9, 0, 1, 8, 4, 7, 2, 3, 11, -1, -1, -1, -1, -1, -1, -1},
    {4, 7, 11, 9, 4, 11, 9, 11, 2, 9, 2, 1, -1, -1, -1, -1},
    {3, 10, 1, 3, 11, 10, 7, 8, 4, -1, -1, -1, -1, -1, -1, -1},
    {1, 11, 10, 1, 4, 11, 1, 0, 4, 7, 11, 4, -1, -1, -1, -1},
    {4, 7, 8, 9, 0, 11, 9, 11, 10, 11, 0, 3, -1, -1, -1, -1},
    {4, 7, 11, 4, 11, 9, 9, 11, 10, -1, -1, -1, -1, -1, -1, -1},
    {9, 5, 4, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {9, 5, 4, 0, 8, 3, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 5, 4, 1, 5, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {8, 5, 4, 8, 3, 5, 3, 1, 5, -1, -1, -1, -1, -1, -1, -1},
    {1, 2, 10, 9, 5, 4, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {3, 0, 8, 1, 2, 10, 4, 9, 5, -1, -1, -1, -1, -1, -1, -1},
    {5, 2, 10, 5, 4, 2, 4, 0, 2, -1, -1, -1, -1, -1, -1, -1},
    {2, 10, 5, 3, 2, 5, 3, 5, 4, 3, 4, 8, -1, -1, -1, -1},
    {9, 5, 4, 2, 3, 11, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 11, 2, 0, 8, 11, 4, 9, 5, -1, -1, -1, -1, -1, -1, -1},
    {0, 5, 4, 0, 1, 5, 2, 3, 11, -1, -1, -1, -1, -1, -1, -1},
    {2, 1, 5, 2, 5, 8, 2, 8, 11, 4, 8, 5, -1, -1, -1, -1},
    {10, 3, 11, 10, 1, 3, 9, 5, 4, -1, -1, -1, -1, -1, -1, -1},
    {4, 9, 5, 0, 8, 1, 8, 10, 1, 8, 11, 10, -1, -1, -1, -1},
    {5, 4, 0, 5, 0, 11, 5, 11, 10, 11, 0, 3, -1, -1, -1, -1},
    {5, 4, 8, 5, 8, 10, 10, 8, 11, -1, -1, -1, -1, -1, -1, -1},
    {9, 7, 8, 5, 7, 9, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {9, 3, 0, 9, 5, 3, 5, 7, 3, -1, -1, -1, -1, -1, -1, -1},
    {0, 7, 8, 0, 1, 7, 1, 5, 7, -1, -1, -1, -1, -1, -1, -1},
    {1, 5, 3, 3, 5, 7, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {9, 7, 8, 9, 5, 7, 10, 1, 2, -1, -1, -1, -1, -1, -1, -1},
    {10, 1, 2, 9, 5, 0, 5, 3, 0, 5, 7, 3, -1, -1, -1, -1},
    {8, 0, 2, 8, 2, 5, 8, 5, 7, 10, 5, 2, -1, -1, -1, -1},
    {2, 10, 5, 2, 5, 3, 3, 5, 7, -1, -1, -1, -1, -1, -1, -1},
    {7, 9, 5, 7, 8, 9, 3, 11, 2, -1, -1, -1, -1, -1, -1, -1},
    {9, 5, 7, 9, 7, 2, 9, 2, 0, 2, 7, 11, -1, -1, -1, -1},
    {2, 3, 11, 0, 1, 8, 1, 7, 8, 1, 5, 7, -1, -1, -1, -1},
    {11, 2, 1, 11, 1, 7, 7, 1, 5, -1, -1, -1, -1, -1, -1, -1},
    {9, 5, 8, 8, 5, 7, 10, 1, 3, 10, 3, 11, -1, -1, -1, -1},
    {5, 7, 0, 5, 0, 9, 7, 11, 0, 1, 0, 10, 11, 10, 0, -1},
    {11, 10, 0, 11, 0, 3, 10, 5, 0, 8, 0, 7, 5, 7, 0, -1},
    {11, 10, 5, 7, 11, 5, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {10, 6, 5, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 8, 3, 5, 10, 6, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {9, 0, 1, 5, 10, 6, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 8, 3, 1, 9, 8, 5, 10, 6, -1, -1, -1, -1, -1, -1, -1},
    {1, 6, 5, 2, 6, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 6, 5, 1, 2, 6, 3, 0, 8, -1, -1, -1, -1, -1, -1, -1},
    {9, 6, 5, 9, 0, 6, 0, 2, 6, -1, -1, -1, -1, -1, -1, -1},
    {5, 9, 8, 5, 8, 2, 5, 2, 6, 3, 2, 8, -1, -1, -1, -1},
    {2, 3, 11, 10, 6, 5, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {11, 0, 8, 11, 2, 0, 10, 6, 5, -1, -1, -1, -1, -1, -1, -1},
    {0, 1, 9, 2, 3, 11, 5, 10, 6, -1, -1, -1, -1, -1, -1, -1},
    {5, 10, 6, 1, 9, 2, 9, 11, 2, 9, 8, 11, -1, -1, -1, -1},
    {6, 3, 11, 6, 5, 3, 5, 1, 3, -1, -1, -1, -1, -1, -1, -1},
    {0, 8, 11, 0, 11, 5, 0, 5, 1, 5, 11, 6, -1, -1, -1, -1},
    {3, 11, 6, 0, 3, 6, 0, 6, 5, 0, 5, 9, -1, -1, -1, -1},
    {6, 5, 9, 6, 9, 11, 11, 9, 8, -1, -1, -1, -1, -1, -1, -1},
    {5, 10, 6, 4, 7, 8, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {4, 3, 0, 4, 7, 3, 6, 5, 10, -1, -1, -1, -1, -1, -1, -1},
    {1, 9, 0, 5, 10, 6, 8, 4, 7, -1, -1, -1, -1, -1, -1, -1},
    {10, 6, 5, 1, 9, 7, 1, 7, 3, 7, 9, 4, -1, -1, -1, -1},
    {6, 1, 2, 6, 5, 1, 4, 7, 8, -1, -1, -1, -1, -1, -1, -1},
    {1, 2, 5, 5, 2, 6, 3, 0, 4, 3, 4, 7, -1, -1, -1, -1},
    {8, 4, 7, 9, 0, 5, 0, 6, 5, 0, 2, 6, -1, -1, -1, -1},
    {7, 3, 9, 7, 9, 4, 3, 2, 9, 5, 9, 6, 2, 6, 9, -1},
    {3, 11, 2, 7, 8, 4, 10, 6, 5, -1, -1, -1, -1, -1, -1, -1},
    {5, 10, 6, 4, 7, 2, 4, 2, 0, 2, 7, 11, -1, -1, -1, -1},
    {0, 1, 9, 4, 7, 8, 2, 3, 11, 5, 10, 6, -1, -1, -1, -1},
    {9, 2, 1, 9, 11, 2, 9, 4, 11, 7, 11, 4, 5, 10, 6, -1},
    {8, 4, 7, 3, 11, 5, 3, 5, 1, 5, 11, 6, -1, -1, -1, -1},
    {5, 1, 11, 5, 11, 6, 1, 0, 11, 7, 11, 4, 0, 4, 11, -1},
    {0, 5, 9, 0, 6, 5, 0, 3, 6, 11, 6, 3, 8, 4, 7, -1},
    {6, 5, 9, 6, 9, 11, 4, 7, 9, 7, 11, 9, -1, -1, -1, -1},
    {10, 4, 9, 6, 4, 10, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {4, 10, 6, 4, 9, 10, 0, 8, 3, -1, -1, -1, -1, -1, -1, -1},
    {10, 0, 1, 10, 6, 0, 6, 4, 0, -1, -1, -1, -1, -1, -1, -1},
    {8, 3, 1, 8, 1, 6, 8, 6, 4, 6, 1, 10, -1, -1, -1, -1},
    {1, 4, 9, 1, 2, 4, 2, 6, 4, -1, -1, -1, -1, -1, -1, -1},
    {3, 0, 8, 1, 2, 9, 2, 4, 9, 2, 6, 4, -1, -1, -1, -1},
    {0, 2, 4, 4, 2, 6, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {8, 3, 2, 8, 2, 4, 4, 2, 6, -1, -1, -1, -1, -1, -1, -1},
    {10, 4, 9, 10, 6, 4, 11, 2, 3, -1, -1, -1, -1, -1, -1, -1},
    {0, 8, 2, 2, 8, 11, 4, 9, 10, 4, 10, 6, -1, -1, -1, -1},
    {3, 11, 2, 0, 1, 6, 0, 6, 4, 6, 1, 10, -1, -1, -1, -1},
    {6, 4, 1, 6, 1, 10, 4, 8, 1, 2, 1, 11, 8, 11, 1, -1},
    {9, 6, 4, 9, 3, 6, 9, 1, 3, 11, 6, 3, -1, -1, -1, -1},
    {8, 11, 1, 8, 1, 0, 11, 6, 1, 9, 1, 4, 6, 4, 1, -1},
    {3, 11, 6, 3, 6, 0, 0, 6, 4, -1, -1, -1, -1, -1, -1, -1},
    {6, 4, 8, 11, 6, 8, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {7, 10, 6, 7, 8, 10, 8, 9, 10, -1, -1, -1, -1, -1, -1, -1},
    {0, 7, 3, 0, 10, 7, 0, 9, 10, 6, 7, 10, -1, -1, -1, -1},
    {10, 6, 7, 1, 10, 7, 1, 7, 8, 1, 8, 0, -1, -1, -1, -1},
    {10, 6, 7, 10, 7, 1, 1, 7, 3, -1, -1, -1, -1, -1, -1, -1},
    {1, 2, 6, 1, 6, 8, 1, 8, 9, 8, 6, 7, -1, -1, -1, -1},
    {2, 6, 9, 2, 9, 1, 6, 7, 9, 0, 9, 3, 7, 3, 9, -1},
    {7, 8, 0, 7, 0, 6, 6, 0, 2, -1, -1, -1, -1, -1, -1, -1},
    {7, 3, 2, 6, 7, 2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {2, 3, 11, 10, 6, 8, 10, 8, 9, 8, 6, 7, -1, -1, -1, -1},
    {2, 0, 7, 2, 7, 11, 0, 9, 7, 6, 7, 10, 9, 10, 7, -1},
    {1, 8, 0, 1, 7, 8, 1, 10, 7, 6, 7, 10, 2, 3, 11, -1},
    {11, 2, 1, 11, 1, 7, 10, 6, 1, 6, 7, 1, -1, -1, -1, -1},
    {8, 9, 6, 8, 6, 7, 9, 1, 6, 11, 6, 3, 1, 3, 6, -1},
    {0, 9, 1, 11, 6, 7, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {7, 8, 0, 7, 0, 6, 3, 11, 0, 11, 6, 0, -1, -1, -1, -1},
    {7, 11, 6, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {7, 6, 11, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {3, 0, 8, 11, 7, 6, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 1, 9, 11, 7, 6, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {8, 1, 9, 8, 3, 1, 11, 7, 6, -1, -1, -1, -1, -1, -1, -1},
    {10, 1, 2, 6, 11, 7, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 2, 10, 3, 0, 8, 6, 11, 7, -1, -1, -1, -1, -1, -1, -1},
    {2, 9, 0, 2, 10, 9, 6, 11, 7, -1, -1, -1, -1, -1, -1, -1},
    {6, 11, 7, 2, 10, 3, 10, 8, 3, 10, 9, 8, -1, -1, -1, -1},
    {7, 2, 3, 6, 2, 7, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {7, 0, 8, 7, 6, 0, 6, 2, 0, -1, -1, -1, -1, -1, -1, -1},
    {2, 7, 6, 2, 3, 7, 0, 1, 9, -1, -1, -1, -1, -1, -1, -1},
    {1, 6, 2, 1, 8, 6, 1, 9, 8, 8, 7, 6, -1, -1, -1, -1},
    {10, 7, 6, 10, 1, 7, 1, 3, 7, -1, -1, -1, -1, -1, -1, -1},
    {10, 7, 6, 1, 7, 10, 1, 8, 7, 1, 0, 8, -1, -1, -1, -1},
    {0, 3, 7, 0, 7, 10, 0, 10, 9, 6, 10, 7, -1, -1, -1, -1},
    {7, 6, 10, 7, 10, 8, 8, 10, 9, -1, -1, -1, -1, -1, -1, -1},
    {6, 8, 4, 11, 8, 6, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {3, 6, 11, 3, 0, 6, 0, 4, 6, -1, -1, -1, -1, -1, -1, -1},
    {8, 6, 11, 8, 4, 6, 9, 0, 1, -1, -1, -1, -1, -1, -1, -1},
    {9, 4, 6, 9, 6, 3, 9, 3, 1, 11, 3, 6, -1, -1, -1, -1},
    {6, 8, 4, 6, 11, 8, 2, 10, 1, -1, -1, -1, -1, -1, -1, -1},
    {1, 2, 10, 3, 0, 11, 0, 6, 11, 0, 4, 6, -1, -1, -1, -1},
    {4, 11, 8, 4, 6, 11, 0, 2, 9, 2, 10, 9, -1, -1, -1, -1},
    {10, 9, 3, 10, 3, 2, 9, 4, 3, 11, 3, 6, 4, 6, 3, -1},
    {8, 2, 3, 8, 4, 2, 4, 6, 2, -1, -1, -1, -1, -1, -1, -1},
    {0, 4, 2, 4, 6, 2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 9, 0, 2, 3, 4, 2, 4, 6, 4, 3, 8, -1, -1, -1, -1},
    {1, 9, 4, 1, 4, 2, 2, 4, 6, -1, -1, -1, -1, -1, -1, -1},
    {8, 1, 3, 8, 6, 1, 8, 4, 6, 6, 10, 1, -1, -1, -1, -1},
    {10, 1, 0, 10, 0, 6, 6, 0, 4, -1, -1, -1, -1, -1, -1, -1},
    {4, 6, 3, 4, 3, 8, 6, 10, 3, 0, 3, 9, 10, 9, 3, -1},
    {10, 9, 4, 6, 10, 4, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {4, 9, 5, 7, 6, 11, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 8, 3, 4, 9, 5, 11, 7, 6, -1, -1, -1, -1, -1, -1, -1},
    {5, 0, 1, 5, 4, 0, 7, 6, 11, -1, -1, -1, -1, -1, -1, -1},
    {11, 7, 6, 8, 3, 4, 3, 5, 4, 3, 1, 5, -1, -1, -1, -1},
    {9, 5, 4, 10, 1, 2, 7, 6, 11, -1, -1, -1, -1, -1, -1, -1},
    {6, 11, 7, 1, 2, 10, 0, 8, 3, 4, 9, 5, -1, -1, -1, -1},
    {7, 6, 11, 5, 4, 10, 4, 2, 10, 4, 0, 2, -1, -1, -1, -1},
    {3, 4, 8, 3, 5, 4, 3, 2, 5, 10, 5, 2, 11, 7, 6, -1},
    {7, 2, 3, 7, 6, 2, 5, 4, 9, -1, -1, -1, -1, -1, -1, -1},
    {9, 5, 4, 0, 8, 6, 0, 6, 2, 6, 8, 7, -1, -1, -1, -1},
    {3, 6, 2, 3, 7, 6, 1, 5, 0, 5, 4, 0, -1, -1, -1, -1},
    {6, 2, 8, 6, 8, 7, 2, 1, 8, 4, 8, 5, 1, 5, 8, -1},
    {9, 5, 4, 10, 1, 6, 1, 7, 6, 1, 3, 7, -1, -1, -1, -1},
    {1, 6, 10, 1, 7, 6, 1, 0, 7, 8, 7, 0, 9, 5, 4, -1},
    {4, 0, 10, 4, 10, 5, 0, 3, 10, 6, 10, 7, 3, 7, 10, -1},
    {7, 6, 10, 7, 10, 8, 5, 4, 10, 4, 8, 10, -1, -1, -1, -1},
    {6, 9, 5, 6, 11, 9, 11, 8, 9, -1, -1, -1, -1, -1, -1, -1},
    {3, 6, 11, 0, 6, 3, 0, 5, 6, 0, 9, 5, -1, -1, -1, -1},
    {0, 11, 8, 0, 5, 11, 0, 1, 5, 5, 6, 11, -1, -1, -1, -1},
    {6, 11, 3, 6, 3, 5, 5, 3, 1, -1, -1, -1, -1, -1, -1, -1},
    {1, 2, 10, 9, 5, 11, 9, 11, 8, 11, 5, 6, -1, -1, -1, -1},
    {0, 11, 3, 0, 6, 11, 0, 9, 6, 5, 6, 9, 1, 2, 10, -1},
    {11, 8, 5, 11, 5, 6, 8, 0, 5, 10, 5, 2, 0, 2, 5, -1},
    {6, 11, 3, 6, 3, 5, 2, 10, 3, 10, 5, 3, -1, -1, -1, -1},
    {5, 8, 9, 5, 2, 8, 5, 6, 2, 3, 8, 2, -1, -1, -1, -1},
    {9, 5, 6, 9, 6, 0, 0, 6, 2, -1, -1, -1, -1, -1, -1, -1},
    {1, 5, 8, 1, 8, 0, 5, 6, 8, 3, 8, 2, 6, 2, 8, -1},
    {1, 5, 6, 2, 1, 6, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 3, 6, 1, 6, 10, 3, 8, 6, 5, 6, 9, 8, 9, 6, -1},
    {10, 1, 0, 10, 0, 6, 9, 5, 0, 5, 6, 0, -1, -1, -1, -1},
    {0, 3, 8, 5, 6, 10, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {10, 5, 6, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {11, 5, 10, 7, 5, 11, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {11, 5, 10, 11, 7, 5, 8, 3, 0, -1, -1, -1, -1, -1, -1, -1},
    {5, 11, 7, 5, 10, 11, 1, 9, 0, -1, -1, -1, -1, -1, -1, -1},
    {10, 7, 5, 10, 11, 7, 9, 8, 1, 8, 3, 1, -1, -1, -1, -1},
    {11, 1, 2, 11, 7, 1, 7, 5, 1, -1, -1, -1, -1, -1, -1, -1},
    {0, 8, 3, 1, 2, 7, 1, 7, 5, 7, 2, 11, -1, -1, -1, -1},
    {9, 7, 5, 9, 2, 7, 9, 0, 2, 2, 11, 7, -1, -1, -1, -1},
    {7, 5, 2, 7, 2, 11, 5, 9, 2, 3, 2, 8, 9, 8, 2, -1},
    {2, 5, 10, 2, 3, 5, 3, 7, 5, -1, -1, -1, -1, -1, -1, -1},
    {8, 2, 0, 8, 5, 2, 8, 7, 5, 10, 2, 5, -1, -1, -1, -1},
    {9, 0, 1, 5, 10, 3, 5, 3, 7, 3, 10, 2, -1, -1, -1, -1},
    {9, 8, 2, 9, 2, 1, 8, 7, 2, 10, 2, 5, 7, 5, 2, -1},
    {1, 3, 5, 3, 7, 5, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 8, 7, 0, 7, 1, 1, 7, 5, -1, -1, -1, -1, -1, -1, -1},
    {9, 0, 3, 9, 3, 5, 5, 3, 7, -1, -1, -1, -1, -1, -1, -1},
    {9, 8, 7, 5, 9, 7, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {5, 8, 4, 5, 10, 8, 10, 11, 8, -1, -1, -1, -1, -1, -1, -1},
    {5, 0, 4, 5, 11, 0, 5, 10, 11, 11, 3, 0, -1, -1, -1, -1},
    {0, 1, 9, 8, 4, 10, 8, 10, 11, 10, 4, 5, -1, -1, -1, -1},
    {10, 11, 4, 10, 4, 5, 11, 3, 4, 9, 4, 1, 3, 1, 4, -1},
    {2, 5, 1, 2, 8, 5, 2, 11, 8, 4, 5, 8, -1, -1, -1, -1},
    {0, 4, 11, 0, 11, 3, 4, 5, 11, 2, 11, 1, 5, 1, 11, -1},
    {0, 2, 5, 0, 5, 9, 2, 11, 5, 4, 5, 8, 11, 8, 5, -1},
    {9, 4, 5, 2, 11, 3, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {2, 5, 10, 3, 5, 2, 3, 4, 5, 3, 8, 4, -1, -1, -1, -1},
    {5, 10, 2, 5, 2, 4, 4, 2, 0, -1, -1, -1, -1, -1, -1, -1},
    {3, 10, 2, 3, 5, 10, 3, 8, 5, 4, 5, 8, 0, 1, 9, -1},
    {5, 10, 2, 5, 2, 4, 1, 9, 2, 9, 4, 2, -1, -1, -1, -1},
    {8, 4, 5, 8, 5, 3, 3, 5, 1, -1, -1, -1, -1, -1, -1, -1},
    {0, 4, 5, 1, 0, 5, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {8, 4, 5, 8, 5, 3, 9, 0, 5, 0, 3, 5, -1, -1, -1, -1},
    {9, 4, 5, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {4, 11, 7, 4, 9, 11, 9, 10, 11, -1, -1, -1, -1, -1, -1, -1},
    {0, 8, 3, 4, 9, 7, 9, 11, 7, 9, 10, 11, -1, -1, -1, -1},
    {1, 10, 11, 1, 11, 4, 1, 4, 0, 7, 4, 11, -1, -1, -1, -1},
    {3, 1, 4, 3, 4, 8, 1, 10, 4, 7, 4, 11, 10, 11, 4, -1},
    {4, 11, 7, 9, 11, 4, 9, 2, 11, 9, 1, 2, -1, -1, -1, -1},
    {9, 7, 4, 9, 11, 7, 9, 1, 11, 2, 11, 1, 0, 8, 3, -1},
    {11, 7, 4, 11, 4, 2, 2, 4, 0, -1, -1, -1, -1, -1, -1, -1},
    {11, 7, 4, 11, 4, 2, 8, 3, 4, 3, 2, 4, -1, -1, -1, -1},
    {2, 9, 10, 2, 7, 9, 2, 3, 7, 7, 4, 9, -1, -1, -1, -1},
    {9, 10, 7, 9, 7, 4, 10, 2, 7, 8, 7, 0, 2, 0, 7, -1},
    {3, 7, 10, 3, 10, 2, 7, 4, 10, 1, 10, 0, 4, 0, 10, -1},
    {1, 10, 2, 8, 7, 4, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {4, 9, 1, 4, 1, 7, 7, 1, 3, -1, -1, -1, -1, -1, -1, -1},
    {4, 9, 1, 4, 1, 7, 0, 8, 1, 8, 7, 1, -1, -1, -1, -1},
    {4, 0, 3, 7, 4, 3, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {4, 8, 7, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {9, 10, 8, 10, 11, 8, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {3, 0, 9, 3, 9, 11, 11, 9, 10, -1, -1, -1, -1, -1, -1, -1},
    {0, 1, 10, 0, 10, 8, 8, 10, 11, -1, -1, -1, -1, -1, -1, -1},
    {3, 1, 10, 11, 3, 10, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 2, 11, 1, 11, 9, 9, 11, 8, -1, -1, -1, -1, -1, -1, -1},
    {3, 0, 9, 3, 9, 11, 1, 2, 9, 2, 11, 9, -1, -1, -1, -1},
    {0, 2, 11, 8, 0, 11, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {3, 2, 11, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {2, 3, 8, 2, 8, 10, 10, 8, 9, -1, -1, -1, -1, -1, -1, -1},
    {9, 10, 2, 0, 9, 2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {2, 3, 8, 2, 8, 10, 0, 1, 8, 1, 10, 8, -1, -1, -1, -1},
    {1, 10, 2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {1, 3, 8, 9, 1, 8, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 9, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {0, 3, 8, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
    {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
};
__constant__ uint8_t kMcNtri[256] = {0, 1, 1, 2, 1, 2, 2, 3, 1, 2, 2, 3, 2, 3, 3, 2, 1, 2, 2, 3, 2, 3, 3, 4, 2, 3, 3, 4, 3, 4, 4, 3, 1, 2, 2, 3, 2, 3, 3, 4, 2, 3, 3, 4, 3, 4, 4, 3, 2, 3, 3, 2, 3, 4, 4, 3, 3, 4, 4, 3, 4, 5, 5, 2, 1, 2, 2, 3, 2, 3, 3, 4, 2, 3, 3, 4, 3, 4, 4, 3, 2, 3, 3, 4, 3, 4, 4, 5, 3, 4, 4, 5, 4, 5, 5, 4, 2, 3, 3, 4, 3, 4, 2, 3, 3, 4, 4, 5, 4, 5, 3, 2, 3, 4, 4, 3, 4, 5, 3, 2, 4, 5, 5, 4, 5, 2, 4, 1, 1, 2, 2, 3, 2, 3, 3, 4, 2, 3, 3, 4, 3, 4, 4, 3, 2, 3, 3, 4, 3, 4, 4, 5, 3, 2, 4, 3, 4, 3, 5, 2, 2, 3, 3, 4, 3, 4, 4, 5, 3, 4, 4, 5, 4, 5, 5, 4, 3, 4, 4, 3, 4, 5, 5, 4, 4, 3, 5, 2, 5, 4, 2, 1, 2, 3, 3, 4, 3, 4, 4, 5, 3, 4, 4, 5, 2, 3, 3, 2, 3, 4, 4, 5, 4, 5, 5, 2, 4, 3, 5, 4, 3, 2, 4, 1, 3, 4, 4, 5, 4, 5, 3, 4, 4, 5, 5, 2, 3, 4, 2, 1, 2, 3, 3, 2, 3, 4, 2, 1, 3, 2, 4, 1, 2, 1, 1, 0};

constexpr int kMcTile = 9;  // a brick and one voxel plane of its +x / +y / +z neighbours

// Key of the vertex on lattice edge (global voxel g, axis a): g biased by
// 2^18 per axis in bits 43-61 / 24-42 / 5-23, the axis in bits 0-4.
__device__ __forceinline__ unsigned long long mc_key(int gx, int gy, int gz, int a)
{
    return ((unsigned long long)(gx + kKeyBias) << 43) | ((unsigned long long)(gy + kKeyBias) << 24) |
           ((unsigned long long)(gz + kKeyBias) << 5) | (unsigned long long)a;
}

__device__ __forceinline__ uint32_t mc_hash(unsigned long long k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (uint32_t)k;
}

// Loads brick b and its +1 planes into tile (s, w), [z][y][x]; a voxel of a
// missing or never-written brick reads (0, 0).  Ends with a barrier.
template <typename T>
__device__ void mc_load_tile(const GridView& g, unsigned int b, float2 (*tile)[kMcTile][kMcTile], int* nb, int4& bc)
{
    __syncthreads();
    if (threadIdx.x == 0)
        bc = g.coords[b];
    __syncthreads();
    if (threadIdx.x < 8) {
        const int c = threadIdx.x;
        int id = c == 0 ? (int)b : find_brick(g, bc.x + (c & 1), bc.y + ((c >> 1) & 1), bc.z + ((c >> 2) & 1));
        if (id >= 0 && (g.coords[id].w & 1))
            id = -1;
        nb[c] = id;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMcTile * kMcTile * kMcTile; i += kThreads) {
        const int x = i % kMcTile, y = (i / kMcTile) % kMcTile, z = i / (kMcTile * kMcTile);
        const int id = nb[(x >> 3) | ((y >> 3) << 1) | ((z >> 3) << 2)];
        float s = 0.f, w = 0.f;
        if (id >= 0)
            ld_pair((const T*)g.payload + (size_t)id * (2 * kVPB), (x & 7) + 8 * (y & 7) + 64 * (z & 7), s, w);
        tile[z][y][x] = make_float2(s, w);
    }
    __syncthreads();
}

// Case of cell `cell` (x + 8y + 64z) from the tile; 0 when a corner is
// unobserved.  f: the corners' sdf.
__device__ __forceinline__ int mc_case(const float2 (*tile)[kMcTile][kMcTile], int cell, float (&f)[8])
{
    const int x = cell & 7, y = (cell >> 3) & 7, z = cell >> 6;
    int ci = 0;
    bool obs = true;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int o = kMcCorner[k];
        const float2 v = tile[z + (o >> 2)][y + ((o >> 1) & 1)][x + (o & 1)];
        f[k] = v.x;
        obs = obs && v.y != 0.0f;
        ci |= (v.x < 0.0f ? 1 : 0) << k;
    }
    return obs ? ci : 0;
}

__device__ __forceinline__ int mc_edge_mask(int ci)
{
    int m = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e)
        m |= (((ci >> kMcEdgeLo[e]) ^ (ci >> kMcEdgeHi[e])) & 1) << e;
    return m;
}

// Inclusive sum over the workgroup (kThreads lanes); *total gets the sum.
__device__ __forceinline__ unsigned int mc_block_scan(unsigned int v, unsigned int* lds, unsigned int* total)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned int u = __shfl_up(v, o);
        if (lane >= o)
            v += u;
    }
    if (lane == 63)
        lds[wv] = v;
    __syncthreads();
    unsigned int before = 0, all = 0;
    for (int k = 0; k < kThreads / 64; ++k) {
        before += k < wv ? lds[k] : 0u;
        all += lds[k];
    }
    __syncthreads();
    *total = all;
    return v + before;
}

// Pass 1: triangles and edge uses of every brick (two cells per thread).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_mc_count(GridView g, unsigned int n, uint2* counts)
{
    __shared__ float2 tile[kMcTile][kMcTile][kMcTile];
    __shared__ int nb[8];
    __shared__ int4 bc;
    __shared__ unsigned int red[2 * (kThreads / 64)];
    for (unsigned int b = blockIdx.x; b < n; b += gridDim.x) {
        mc_load_tile<T>(g, b, tile, nb, bc);
        unsigned int nt = 0, ne = 0;
        for (int cell = 2 * threadIdx.x; cell < 2 * threadIdx.x + 2; ++cell) {
            float f[8];
            const int ci = mc_case(tile, cell, f);
            nt += kMcNtri[ci];
            ne += ci == 0 || ci == 255 ? 0u : (unsigned int)__popc(mc_edge_mask(ci));
        }
        unsigned int tt, te;
        mc_block_scan(nt, red, &tt);
        mc_block_scan(ne, red + kThreads / 64, &te);
        if (threadIdx.x == 0)
            counts[b] = make_uint2(tt, te);
    }
}

// Pass 2: the triangles' vertex keys, at brick offset + cell order, and
// every used edge's key into the weld table (open addressing, key 0 empty).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_mc_emit(GridView g, unsigned int n, const unsigned long long* tri_off,
                                                      unsigned long long* tkeys, unsigned long long* hkeys,
                                                      uint32_t hmask)
{
    __shared__ float2 tile[kMcTile][kMcTile][kMcTile];
    __shared__ int nb[8];
    __shared__ int4 bc;
    __shared__ unsigned int red[kThreads / 64];
    for (unsigned int b = blockIdx.x; b < n; b += gridDim.x) {
        mc_load_tile<T>(g, b, tile, nb, bc);
        int cis[2];
        float f[8];
        unsigned int nt = 0;
        for (int k = 0; k < 2; ++k) {
            cis[k] = mc_case(tile, 2 * threadIdx.x + k, f);
            nt += kMcNtri[cis[k]];
        }
        unsigned int tot;
        const unsigned int incl = mc_block_scan(nt, red, &tot);
        unsigned long long t = tri_off[b] + (incl - nt);
        for (int k = 0; k < 2; ++k) {
            const int ci = cis[k];
            if (ci == 0 || ci == 255)
                continue;
            const int cell = 2 * threadIdx.x + k;
            const int gx = bc.x * 8 + (cell & 7), gy = bc.y * 8 + ((cell >> 3) & 7), gz = bc.z * 8 + (cell >> 6);
            unsigned long long ek[12];
            const int em = mc_edge_mask(ci);
#pragma unroll
            for (int e = 0; e < 12; ++e) {
                const int o = kMcCorner[kMcEdgeLo[e]];
                ek[e] = mc_key(gx + (o & 1), gy + ((o >> 1) & 1), gz + (o >> 2), kMcEdgeAxis[e]);
                if (!((em >> e) & 1))
                    continue;
                uint32_t h = mc_hash(ek[e]) & hmask;
                for (uint32_t p = 0; p <= hmask; ++p) {
                    const unsigned long long cur = hkeys[h];
                    if (cur == ek[e])
                        break;
                    if (cur == 0ull) {
                        const unsigned long long prev = atomicCAS(&hkeys[h], 0ull, ek[e]);
                        if (prev == 0ull || prev == ek[e])
                            break;
                    }
                    h = (h + 1) & hmask;
                }
            }
            const int m = kMcNtri[ci];
            for (int q = 0; q < m; ++q, ++t) {
                tkeys[3 * t + 0] = ek[kMcTri[ci][3 * q + 0]];
                tkeys[3 * t + 1] = ek[kMcTri[ci][3 * q + 2]];
                tkeys[3 * t + 2] = ek[kMcTri[ci][3 * q + 1]];
            }
        }
    }
}

// Pass 3: a vertex id for every occupied slot of the weld table.  A
// workgroup numbers a run of kMcIdsPer slots with one atomic on the
// counter (a counter taken by every wavefront would queue the atomics of
// the whole chip on one address).
constexpr int kMcIdsPer = 16;  // slots per thread and run
__global__ __launch_bounds__(kThreads) void k_mc_ids(const unsigned long long* hkeys, uint32_t* hids, size_t cap,
                                                     unsigned long long* counter)
{
    __shared__ unsigned int red[kThreads / 64];
    __shared__ unsigned long long wg_base;
    const size_t run = (size_t)kThreads * kMcIdsPer;
    for (size_t base = (size_t)blockIdx.x * run; base < cap; base += (size_t)gridDim.x * run) {
        unsigned int mine = 0;  // occupied slots of this thread: base + threadIdx.x + kThreads k
        for (int k = 0; k < kMcIdsPer; ++k) {
            const size_t i = base + threadIdx.x + (size_t)kThreads * k;
            mine |= (i < cap && hkeys[i] != 0ull ? 1u : 0u) << k;
        }
        // ids in thread order within the run (vertex order is free)
        unsigned int total;
        const unsigned int cnt = (unsigned int)__popc(mine);
        const unsigned int before = mc_block_scan(cnt, red, &total) - cnt;
        if (threadIdx.x == 0)
            wg_base = total ? atomicAdd(counter, (unsigned long long)total) : 0ull;
        __syncthreads();
        unsigned long long next = wg_base + before;
        for (int k = 0; k < kMcIdsPer; ++k) {
            const size_t i = base + threadIdx.x + (size_t)kThreads * k;
            if ((mine >> k) & 1u)
                hids[i] = (uint32_t)(next++);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int mc_voxel(const GridView& g, int gx, int gy, int gz, int& v)
{
    v = (gx & 7) + 8 * (gy & 7) + 64 * (gz & 7);
    return find_brick(g, gx >> 3, gy >> 3, gz >> 3);
}

// Pass 4: position, colour and key of every vertex, from its edge.
template <typename T, bool kColor>
__global__ __launch_bounds__(kThreads) void k_mc_vertices(GridView g, const unsigned long long* hkeys,
                                                          const uint32_t* hids, size_t cap, double* verts,
                                                          double* cols, unsigned long long* vkeys)
{
    const double vl = (double)g.vl, half = vl * 0.5;
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < cap; i += stride) {
        const unsigned long long key = hkeys[i];
        if (key == 0ull)
            continue;
        const uint32_t id = hids[i];
        const int a = (int)(key & 31ull);
        const int gx = (int)((key >> 43) & 0x7FFFFull) - kKeyBias;
        const int gy = (int)((key >> 24) & 0x7FFFFull) - kKeyBias;
        const int gz = (int)((key >> 5) & 0x7FFFFull) - kKeyBias;
        int v0, v1;
        const int b0 = mc_voxel(g, gx, gy, gz, v0);
        const int b1 = mc_voxel(g, gx + (a == 0), gy + (a == 1), gz + (a == 2), v1);
        float s0 = 0.f, s1 = 0.f, w;
        if (b0 >= 0)
            ld_pair((const T*)g.payload + (size_t)b0 * (2 * kVPB), v0, s0, w);
        if (b1 >= 0)
            ld_pair((const T*)g.payload + (size_t)b1 * (2 * kVPB), v1, s1, w);
        const double f0 = fabs((double)s0), f1 = fabs((double)s1);
        double p[3] = {half + vl * (double)gx, half + vl * (double)gy, half + vl * (double)gz};
        p[a] += f0 * vl / (f0 + f1);
        for (int k = 0; k < 3; ++k)
            verts[3 * (size_t)id + k] = p[k];
        if (kColor) {
            for (int k = 0; k < 3; ++k) {
                const double c0 = b0 >= 0 ? (double)g.rgb[((size_t)b0 * 3 + k) * kVPB + v0] / 255.0 : 0.0;
                const double c1 = b1 >= 0 ? (double)g.rgb[((size_t)b1 * 3 + k) * kVPB + v1] / 255.0 : 0.0;
                cols[3 * (size_t)id + k] = (f1 * c0 + f0 * c1) / (f0 + f1);
            }
        }
        if (vkeys)
            vkeys[id] = key;
    }
}

__device__ __forceinline__ uint32_t mc_lookup(const unsigned long long* hkeys, const uint32_t* hids, uint32_t hmask,
                                              unsigned long long key)
{
    uint32_t h = mc_hash(key) & hmask;
    for (uint32_t p = 0; p <= hmask; ++p) {
        const unsigned long long cur = hkeys[h];
        if (cur == key)
            return hids[h];
        if (cur == 0ull)
            break;
        h = (h + 1) & hmask;
    }
    return 0xFFFFFFFFu;  // unreachable: every vertex key was inserted in pass 2
}

// Pass 5: vertex indices of the triangles.
__global__ __launch_bounds__(kThreads) void k_mc_faces(const unsigned long long* tkeys, size_t n3,
                                                       const unsigned long long* hkeys, const uint32_t* hids,
                                                       uint32_t hmask, int32_t* faces)
{
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n3; i += stride)
        faces[i] = (int32_t)mc_lookup(hkeys, hids, hmask, tkeys[i]);
}

// Pass 6: every triangle adds its unnormalised normal to its vertices.
__global__ __launch_bounds__(kThreads) void k_mc_face_normals(const double* verts, const int32_t* faces, size_t nt,
                                                              uint32_t nv, double* vn)
{
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < nt; t += stride) {
        const int32_t i0 = faces[3 * t], i1 = faces[3 * t + 1], i2 = faces[3 * t + 2];
        if ((uint32_t)i0 >= nv || (uint32_t)i1 >= nv || (uint32_t)i2 >= nv)
            continue;  // unreachable: every index comes from the weld table
        double a[3], b[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = verts[3 * (size_t)i1 + k] - verts[3 * (size_t)i0 + k];
            b[k] = verts[3 * (size_t)i2 + k] - verts[3 * (size_t)i0 + k];
        }
        const double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        // the hardware fp64 add (device memory of this process; the order of
        // the additions is free anyway)
        for (int k = 0; k < 3; ++k) {
            unsafeAtomicAdd(&vn[3 * (size_t)i0 + k], n[k]);
            unsafeAtomicAdd(&vn[3 * (size_t)i1 + k], n[k]);
            unsafeAtomicAdd(&vn[3 * (size_t)i2 + k], n[k]);
        }
    }
}

// Pass 7: unit normals (Eigen's normalize: a zero vector stays zero).
__global__ __launch_bounds__(kThreads) void k_mc_normalize(double* vn, size_t nv)
{
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < nv; i += stride) {
        double* p = vn + 3 * i;
        const double z = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        if (z > 0.0) {
            const double r = sqrt(z);
            p[0] /= r;
            p[1] /= r;
            p[2] /= r;
        }
    }
}

}  // namespace

extern "C" {

tsdf_status tsdf_marching_cubes(tsdf_grid* g, double* vertices, double* vertex_colors, double* vertex_normals,
                                int32_t* triangles, uint64_t* vertex_keys, uint64_t vertex_cap, uint64_t triangle_cap,
                                uint64_t* n_vertices, uint64_t* n_triangles, void* stream)
{
    if (!g || !n_vertices || !n_triangles)
        return fail(TSDF_BAD_ARG, "NULL grid or count");
    if ((vertex_cap && !vertices) || (triangle_cap && !triangles))
        return fail(TSDF_BAD_ARG, "NULL vertex or triangle buffer with a nonzero capacity");
    if (vertex_colors && !g->rgb)
        return fail(TSDF_BAD_ARG, "vertex colours asked of a grid without colour");
    tsdf_stats_out st;
    tsdf_status r = tsdf_stats(g, &st, stream);
    if (r != TSDF_OK)
        return r;
    *n_vertices = 0;
    *n_triangles = 0;
    const unsigned int n = (unsigned int)st.n_bricks;
    if (n == 0)
        return TSDF_OK;
    {
        // every block ever inserted lies in the control block's bounding box
        const Ctrl& c = *g->ctrl_host;
        for (int a = 0; a < 3; ++a)
            if (c.bbox_min[a] < -(kKeyBias / 8) || c.bbox_max[a] >= kKeyBias / 8 - 1)
                return fail(TSDF_BAD_ARG, "map extent beyond the vertex key range (+-32767 blocks per axis)");
    }
    DeviceScope ds(g->p.device);
    hipStream_t s = (hipStream_t)stream;
    const bool f16 = g->p.storage == TSDF_STORAGE_F16;
    const int bgrid = (int)(n < 65536u ? n : 65536u);
    const int lgrid = g->n_cu * 8;
    uint2* counts = nullptr;
    unsigned long long *tri_off = nullptr, *tkeys = nullptr, *hkeys = nullptr, *counter = nullptr;
    uint32_t* hids = nullptr;
    tsdf_status out = TSDF_OK;
    auto check = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && out == TSDF_OK)
            out = fail(TSDF_HIP_ERROR, std::string("marching cubes: ") + what + ": " + hipGetErrorString(e));
        return out == TSDF_OK;
    };
    std::vector<uint2> hc(n);
    std::vector<unsigned long long> off(n);
    unsigned long long T = 0, E = 0, V = 0;
    if (check(hipMalloc(&counts, (size_t)n * sizeof(uint2)), "hipMalloc")) {
        if (f16)
            hipLaunchKernelGGL(k_mc_count<__half>, dim3(bgrid), dim3(kThreads), 0, s, view(g), n, counts);
        else
            hipLaunchKernelGGL(k_mc_count<float>, dim3(bgrid), dim3(kThreads), 0, s, view(g), n, counts);
        check(hipGetLastError(), "count");
        check(hipMemcpyAsync(hc.data(), counts, (size_t)n * sizeof(uint2), hipMemcpyDeviceToHost, s), "copy counts");
        check(hipStreamSynchronize(s), "synchronize");
    }
    if (out == TSDF_OK) {
        for (unsigned int b = 0; b < n; ++b) {
            off[b] = T;
            T += hc[b].x;
            E += hc[b].y;
        }
    }
    size_t hcap = 1024;
    while (out == TSDF_OK && hcap < 2 * E)
        hcap <<= 1;
    if (out == TSDF_OK && T > 0 && hcap > (size_t)1 << 32)
        out = fail(TSDF_BAD_ARG, "mesh too large for 32-bit vertex indices");
    if (out == TSDF_OK && T > 0 && check(hipMalloc(&tri_off, (size_t)n * sizeof(unsigned long long)), "hipMalloc") &&
        check(hipMalloc(&tkeys, (size_t)T * 3 * sizeof(unsigned long long)), "hipMalloc") &&
        check(hipMalloc(&hkeys, hcap * sizeof(unsigned long long)), "hipMalloc") &&
        check(hipMalloc(&hids, hcap * sizeof(uint32_t)), "hipMalloc") &&
        check(hipMalloc(&counter, sizeof(unsigned long long)), "hipMalloc") &&
        check(hipMemcpyAsync(tri_off, off.data(), (size_t)n * sizeof(unsigned long long), hipMemcpyHostToDevice, s),
              "copy offsets")) {
        hipLaunchKernelGGL(k_fill_u64, dim3(lgrid), dim3(kThreads), 0, s, hkeys, hcap);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, s, counter, (size_t)1);
        const uint32_t hmask = (uint32_t)(hcap - 1);
        if (f16)
            hipLaunchKernelGGL(k_mc_emit<__half>, dim3(bgrid), dim3(kThreads), 0, s, view(g), n, tri_off, tkeys, hkeys,
                               hmask);
        else
            hipLaunchKernelGGL(k_mc_emit<float>, dim3(bgrid), dim3(kThreads), 0, s, view(g), n, tri_off, tkeys, hkeys,
                               hmask);
        hipLaunchKernelGGL(k_mc_ids, dim3(lgrid), dim3(kThreads), 0, s, hkeys, hids, hcap, counter);
        check(hipGetLastError(), "emit");
        check(hipMemcpyAsync(&V, counter, sizeof(V), hipMemcpyDeviceToHost, s), "copy vertex count");
        check(hipStreamSynchronize(s), "synchronize");
        if (out == TSDF_OK && V <= vertex_cap && T <= triangle_cap) {
            const bool col = vertex_colors != nullptr;
            if (f16 && col)
                hipLaunchKernelGGL((k_mc_vertices<__half, true>), dim3(lgrid), dim3(kThreads), 0, s, view(g), hkeys,
                                   hids, hcap, vertices, vertex_colors, (unsigned long long*)vertex_keys);
            else if (f16)
                hipLaunchKernelGGL((k_mc_vertices<__half, false>), dim3(lgrid), dim3(kThreads), 0, s, view(g), hkeys,
                                   hids, hcap, vertices, vertex_colors, (unsigned long long*)vertex_keys);
            else if (col)
                hipLaunchKernelGGL((k_mc_vertices<float, true>), dim3(lgrid), dim3(kThreads), 0, s, view(g), hkeys,
                                   hids, hcap, vertices, vertex_colors, (unsigned long long*)vertex_keys);
            else
                hipLaunchKernelGGL((k_mc_vertices<float, false>), dim3(lgrid), dim3(kThreads), 0, s, view(g), hkeys,
                                   hids, hcap, vertices, vertex_colors, (unsigned long long*)vertex_keys);
            hipLaunchKernelGGL(k_mc_faces, dim3(lgrid), dim3(kThreads), 0, s, tkeys, (size_t)T * 3, hkeys, hids, hmask,
                               triangles);
            if (vertex_normals) {
                hipLaunchKernelGGL(k_fill_f64, dim3(lgrid), dim3(kThreads), 0, s, vertex_normals, (size_t)V * 3);
                hipLaunchKernelGGL(k_mc_face_normals, dim3(lgrid), dim3(kThreads), 0, s, vertices, triangles,
                                   (size_t)T, (uint32_t)V, vertex_normals);
                hipLaunchKernelGGL(k_mc_normalize, dim3(lgrid), dim3(kThreads), 0, s, vertex_normals, (size_t)V);
            }
            check(hipGetLastError(), "vertices");
            check(hipStreamSynchronize(s), "synchronize");
        }
    }
    (void)hipFree(counts);
    (void)hipFree(tri_off);
    (void)hipFree(tkeys);
    (void)hipFree(hkeys);
    (void)hipFree(hids);
    (void)hipFree(counter);
    if (out != TSDF_OK)
        return out;
    *n_vertices = V;
    *n_triangles = T;
    return TSDF_OK;
}

}  // extern "C"
