/*
 * ref_nanovdb.c — test-only shim that compiles the reference's own NanoVDB
 * reader (extern/nanovdb/PNanoVDB.h, PNANOVDB_C mode, 32.3.3) as it lies in
 * the reference tree, so the tests can read grids exactly as the reference's
 * medium code does (src/Shaders/materials/medium.hlsli:85-101, 130-145).
 * Built by oracle/Makefile into oracle/_ref/ only when the reference is present.
 */
#define PNANOVDB_C
#define PNANOVDB_BUF_BOUNDS_CHECK
#include "PNanoVDB.h"

static pnanovdb_buf_t mk(const void* data, uint64_t bytes) { return pnanovdb_make_buf((uint32_t*)data, bytes / 4); }

/* readGrid: value at floor(pos) through a read accessor (medium.hlsli:85-88) */
float ref_read_float(const void* data, uint64_t bytes, int i, int j, int k) {
    pnanovdb_buf_t buf = mk(data, bytes);
    pnanovdb_grid_handle_t grid = {pnanovdb_address_null()};
    pnanovdb_root_handle_t root = pnanovdb_tree_get_root(buf, pnanovdb_grid_get_tree(buf, grid));
    pnanovdb_readaccessor_t acc;
    pnanovdb_readaccessor_init(&acc, root);
    pnanovdb_coord_t ijk = {i, j, k};
    pnanovdb_address_t a = pnanovdb_readaccessor_get_value_address(PNANOVDB_GRID_TYPE_FLOAT, buf, &acc, &ijk);
    return pnanovdb_read_float(buf, a);
}

/* root maximum (majorant source, medium.hlsli:137) and index bbox (intersection.hlsli:103-107) */
void ref_root_info(const void* data, uint64_t bytes, float* max_value, int bbox[6]) {
    pnanovdb_buf_t buf = mk(data, bytes);
    pnanovdb_grid_handle_t grid = {pnanovdb_address_null()};
    pnanovdb_root_handle_t root = pnanovdb_tree_get_root(buf, pnanovdb_grid_get_tree(buf, grid));
    *max_value = pnanovdb_read_float(buf, pnanovdb_root_get_max_address(PNANOVDB_GRID_TYPE_FLOAT, buf, root));
    pnanovdb_coord_t mn = pnanovdb_root_get_bbox_min(buf, root), mx = pnanovdb_root_get_bbox_max(buf, root);
    bbox[0] = mn.x; bbox[1] = mn.y; bbox[2] = mn.z; bbox[3] = mx.x; bbox[4] = mx.y; bbox[5] = mx.z;
}

/* world <-> index maps (PNanoVDB.h:2036-2058) */
void ref_world_to_index(const void* data, uint64_t bytes, const float* p, int dir, float* out) {
    pnanovdb_buf_t buf = mk(data, bytes);
    pnanovdb_grid_handle_t grid = {pnanovdb_address_null()};
    pnanovdb_vec3_t v = {p[0], p[1], p[2]};
    pnanovdb_vec3_t r = dir ? pnanovdb_grid_world_to_index_dirf(buf, grid, &v) : pnanovdb_grid_world_to_indexf(buf, grid, &v);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
