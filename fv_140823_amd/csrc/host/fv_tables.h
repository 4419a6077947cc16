/*
 * Hexahedron element tables shared by the host forest and the flattener.
 * Node index = i + 2j + 4k (x fastest); child index = corner index.
 * Face node lists follow fcnd2elnd (operatetp.c:117-147), opposite faces fcopp
 * (operatetp.c:485-492).  Faces: 0 -y, 1 +x, 2 +y, 3 -x, 4 -z, 5 +z.
 */
#ifndef FV_TABLES_H
#define FV_TABLES_H

static const int FV_FACE_NODES[6][4] = {
  {0, 1, 5, 4}, {3, 7, 5, 1}, {2, 6, 7, 3}, {2, 0, 4, 6}, {2, 3, 1, 0}, {6, 4, 5, 7}};
static const int FV_FACE_OPP[6] = {2, 3, 0, 1, 5, 4};
static const int FV_FACE_AXIS[6] = {1, 0, 1, 0, 2, 2};
static const int FV_FACE_SIGN[6] = {-1, 1, 1, -1, -1, 1};
/* side index in the reference's bc order -x,+x,-y,+y,-z,+z (inputset.c:99-104) */
static const int FV_FACE_SIDE[6] = {2, 1, 3, 0, 4, 5};
/* volume tetrahedra of a hexahedron (meshproperties.c:79-85) */
static const int FV_VOL_TETS[6][3] = {
  {1, 3, 5}, {5, 3, 7}, {5, 7, 4}, {3, 2, 7}, {7, 2, 6}, {7, 6, 4}};

/* face classes of the flattened mesh */
enum { FV_CLS_SAME = 0, FV_CLS_FINER = 1, FV_CLS_COARSER = 2, FV_CLS_BOUNDARY = 3 };

#endif
