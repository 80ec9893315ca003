/*
 * kmc.h -- C ABI of the MI355X marching-cubes isosurface library (libkmc.so).
 *
 * Drop-in boundary for the isosurface path of the Kyoto Visualization System:
 * kvs::MarchingCubes / kvs::Isosurface over kvs::StructuredVolumeObject
 * (Source/Core/Visualization/Mapper/MarchingCubes.h:22-63, Isosurface.h:22-55).
 * The C++ mapper keeps the reference's constructor and exec() surface and calls
 * these entry points; everything below takes plain pointers and sizes.
 *
 * Results are bit-identical to the reference: the same triangles in the same
 * order with the same winding, the same vertex numbering, and coordinates and
 * normals rounded exactly as the reference rounds them.
 */
#ifndef KMC_H
#define KMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KMC_MAX_LEVELS_PER_PASS 8

/* Status codes.  The first four mirror the reference's error exits, which print
 * kvsMessageError, leave isSuccess() false and return nullptr. */
typedef enum kmc_status {
    KMC_OK = 0,
    KMC_ERR_NULL_INPUT = 1,       /* "Input object is NULL."               MarchingCubes.cpp:50-54  */
    KMC_ERR_NOT_STRUCTURED = 2,   /* "Input object is not structured..."   MarchingCubes.cpp:56-61  */
    KMC_ERR_NOT_SCALAR = 3,       /* "Input volume is not sclar field data." MarchingCubes.cpp:63-67 */
    KMC_ERR_UNSUPPORTED_TYPE = 4, /* "Unsupported data type '%s'."         MarchingCubes.cpp:125-129 */
    KMC_ERR_INVALID_ARGUMENT = 5,
    KMC_ERR_OUT_OF_MEMORY = 6,    /* "Cannot allocate memory for the vertex map." MarchingCubes.cpp:257-261 */
    KMC_ERR_DEVICE = 7,           /* a HIP runtime call failed */
    KMC_ERR_BUDGET = 8,           /* serving: request does not fit the HBM budget */
    KMC_ERR_NOT_READY = 9,        /* serving: result not finished yet */
    KMC_ERR_NO_DEVICE = 10,       /* no HIP device visible */
    KMC_ERR_SCRATCH = 11,         /* kmc_extract_noalloc: work exceeds KMC_NOALLOC_SCRATCH_BYTES */
    KMC_ERR_CAPACITY = 12,        /* kmc_extract_noalloc: output arrays too small (sizes reported) */
    KMC_ERR_CELL_TYPE = 13,       /* unstructured: not tetrahedra / hexahedra / pyramid / prism cells */
    KMC_ERR_FILE = 14,            /* a file cannot be read or written */
    KMC_ERR_NOT_TRIANGLES = 15    /* "Input object is not triangle polygon." (PolygonExporter.cpp:80-85) */
} kmc_status;

/* Value types: the order of kvs::Type::TypeID (Source/Core/Utility/Type.h:40-53). */
typedef enum kmc_type {
    KMC_INT8 = 0, KMC_INT16 = 1, KMC_INT32 = 2, KMC_INT64 = 3,
    KMC_UINT8 = 4, KMC_UINT16 = 5, KMC_UINT32 = 6, KMC_UINT64 = 7,
    KMC_REAL32 = 8, KMC_REAL64 = 9
} kmc_type;

/* kvs::PolygonObject::NormalType (Source/Core/Visualization/Object/PolygonObject.h:50-55). */
typedef enum kmc_normal_type { KMC_VERTEX_NORMAL = 0, KMC_POLYGON_NORMAL = 1 } kmc_normal_type;

/* Where a buffer lives. */
typedef enum kmc_memory { KMC_MEMORY_HOST = 0, KMC_MEMORY_DEVICE = 1 } kmc_memory;

/* Volume kind: only structured volumes are accepted (the check of
 * StructuredVolumeObject::DownCast at MarchingCubes.cpp:56-61). */
typedef enum kmc_volume_kind { KMC_STRUCTURED = 0, KMC_UNSTRUCTURED = 1 } kmc_volume_kind;

/* The fields of kvs::StructuredVolumeObject the path reads
 * (StructuredVolumeObject.h:23-99, VolumeObjectBase.h:34-53, ObjectBase.h:44-47).
 * Rectilinear and curvilinear grids are treated, as by the reference, as a
 * uniform grid between min_coord and max_coord. */
typedef struct kmc_volume {
    const void* values;          /* resolution[0]*resolution[1]*resolution[2]*veclen values, x fastest */
    int32_t memory;              /* kmc_memory of values */
    int32_t type;                /* kmc_type */
    uint32_t resolution[3];
    uint32_t veclen;             /* must be 1 */
    int32_t kind;                /* kmc_volume_kind */
    int32_t has_min_max_values;  /* else computed as GetMinMaxValues<T> (StructuredVolumeObject.cpp:66-111) */
    double min_value, max_value;
    int32_t has_min_max_coords;  /* else (0,0,0) .. resolution-1 (StructuredVolumeObject.cpp:355-363) */
    float min_coord[3], max_coord[3];
} kmc_volume;

/* One extracted surface: the arrays of the kvs::PolygonObject the reference
 * returns (GeometryObjectBase.h:44-46, PolygonObject.h:58-61), AoS x,y,z. */
typedef struct kmc_surface {
    double isolevel;
    uint64_t n_vertices;         /* coordinate triples; 3*n_triangles when duplicated */
    uint64_t n_triangles;
    uint64_t n_connections;      /* uint32 entries, 3*n_triangles; 0 when duplicated */
    uint64_t n_normals;          /* normal triples: per vertex (VertexNormal) or per triangle */
    float* coords;
    uint32_t* connections;
    float* normals;
    int32_t memory;              /* kmc_memory of the three arrays */
    int32_t polygon_type;        /* 3 = Triangle */
    int32_t normal_type;         /* kmc_normal_type as set on the reference's output */
    int32_t color_type;          /* 1 = PolygonColor */
    int32_t empty_range;         /* 1 when Math::Equal(min,max) skipped extraction (MarchingCubes.cpp:108-110) */
    double min_value, max_value; /* the volume min/max used (clamp range and transfer-function range) */
    float min_coord[3], max_coord[3]; /* min/max object coordinates copied to the polygon (MapperBase.cpp:116-121) */
    void* internal;              /* ownership handle; release with kmc_surface_release */
    uint8_t* colors;             /* per-vertex RGB (kmc_slice_plane: color_type 0 = VertexColor); else NULL */
    uint64_t n_colors;           /* RGB triples */
} kmc_surface;

/* Select the HIP device used by this thread's calls (default 0). */
kmc_status kmc_init(int device);

/* Human-readable text of a status, matching the reference's messages. */
const char* kmc_status_string(kmc_status status);

/*
 * Plain extraction at one isolevel: the drop-in for
 * kvs::MarchingCubes(volume, isolevel, normal_type, duplication, tf)
 * (MarchingCubes.cpp:25-37, exec :46-79, mapping :87-132).  VertexNormal forces
 * duplication off (:69-73).  The outputs are allocated by the library in
 * out_memory and belong to *out until kmc_surface_release.  stream is a
 * hipStream_t (NULL: the null stream); the call returns after the results are
 * complete.
 */
kmc_status kmc_extract(const kmc_volume* volume, double isolevel, int32_t normal_type,
                       int32_t duplication, int32_t out_memory, void* stream, kmc_surface* out);

/*
 * Several isolevels of one volume, classified in the same pass over the
 * volume (the volume is read once per KMC_MAX_LEVELS_PER_PASS levels).
 * outs[i] is identical to kmc_extract at isolevels[i]; there is no reference
 * counterpart beyond calling kvs::MarchingCubes once per level.
 */
kmc_status kmc_extract_levels(const kmc_volume* volume, const double* isolevels, uint32_t n_levels,
                              int32_t normal_type, int32_t duplication, int32_t out_memory,
                              void* stream, kmc_surface* outs);

/*
 * kmc_extract without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * The volume must be in device memory; the caller passes device output arrays
 * with their capacities in elements (floats for coords and normals, uint32
 * for connections; connections may be NULL with duplication).  Every work
 * array lives in one fixed device array of KMC_NOALLOC_SCRATCH_BYTES that the
 * library holds on each GPU it runs on until the process ends:
 *   - KMC_ERR_SCRATCH when the volume's work (about 48 bytes per 64-node chunk
 *     and 24 per node row) or the emit's records (about 8 bytes per vertex and
 *     9 per triangle, 28 more per vertex with VertexNormal) do not fit;
 *   - KMC_ERR_CAPACITY when the outputs do not fit: *out then holds the sizes
 *     (n_vertices, n_triangles, n_connections, n_normals) and nothing was written.
 * On KMC_OK *out describes the caller's arrays (release is not needed).
 * Calls on one GPU share its scratch array: the library runs them one at a
 * time (a per-device lock); kernels of calls made from other processes on the
 * same GPU use their own process's array.  The call returns when the surface
 * is complete.
 */
#define KMC_NOALLOC_SCRATCH_BYTES (256ull << 20)

kmc_status kmc_extract_noalloc(const kmc_volume* volume, double isolevel, int32_t normal_type,
                               int32_t duplication, void* stream, float* coords, uint64_t coords_capacity,
                               uint32_t* connections, uint64_t connections_capacity, float* normals,
                               uint64_t normals_capacity, kmc_surface* out);

/*
 * kvs::SlicePlane over a structured volume: the drop-in for
 * SlicePlane(volume, coefficients, transfer_function) (SlicePlane.cpp:26-36,
 * exec :65-85, mapping :93-150, extract_plane<T> :165-280).  coefficients:
 * plane(x,y,z) = c[0]*x + c[1]*y + c[2]*z + c[3] over node indices (the
 * point / normal form stores { normal, -point.dot(normal) }).  The surface
 * holds, like the reference's polygon: coords per triangle corner
 * (n_vertices = 3 * n_triangles), one normal per triangle
 * (-(v2 - v0) x (v1 - v0)), one colour per corner (colors, n_colors) from
 * the colour map color_table (3 bytes per entry, color_resolution entries:
 * ColorMap::table()) over color_range ([min, max]; NULL: the range
 * MapperBase::setRange gives a transfer function without one, MapperBase.cpp:54-83);
 * polygon_type 3, color_type 0 (VertexColor), normal_type PolygonNormal.
 * Unstructured volumes: kmc_slice_plane_unstructured.
 */
kmc_status kmc_slice_plane(const kmc_volume* volume, const float coefficients[4], const uint8_t* color_table,
                           uint32_t color_resolution, const float* color_range, int32_t out_memory, void* stream,
                           kmc_surface* out);

/* ------------------------------------------------------------------------- */
/* Unstructured volumes: the unstructured branch of kvs::Isosurface          */
/* (Isosurface.cpp:192-257) and kvs::SlicePlane over unstructured cells.      */
/* ------------------------------------------------------------------------- */

/* kvs::UnstructuredVolumeObject::CellType (UnstructuredVolumeObject.h:31-41)
 * of the cell types with a marching table. */
typedef enum kmc_cell_type {
    KMC_CELL_TETRAHEDRA = 1, KMC_CELL_HEXAHEDRA = 2, KMC_CELL_PYRAMID = 5, KMC_CELL_PRISM = 7
} kmc_cell_type;

/* The fields of kvs::UnstructuredVolumeObject the path reads
 * (UnstructuredVolumeObject.h:28-99): node values, node coordinates (3 floats
 * per node) and the cells' node indices (4 / 8 / 5 / 6 per cell), all in the
 * same memory. */
typedef struct kmc_unstructured_volume {
    const void* values;          /* n_nodes values */
    const float* coords;         /* 3 * n_nodes */
    const uint32_t* connections; /* nodes per cell * n_cells */
    int32_t memory;              /* kmc_memory of the three arrays */
    int32_t type;                /* kmc_type of values */
    uint64_t n_nodes, n_cells;
    int32_t cell_type;           /* kmc_cell_type */
    uint32_t veclen;             /* must be 1 */
    int32_t has_min_max_values;  /* else GetMinMaxValues in connection order (UnstructuredVolumeObject.cpp:123-141) */
    double min_value, max_value;
    int32_t has_min_max_coords;  /* else updateMinMaxCoords in connection order (UnstructuredVolumeObject.cpp:425-455) */
    float min_coord[3], max_coord[3];
} kmc_unstructured_volume;

/*
 * kvs::Isosurface over an unstructured volume: MarchingTetrahedra /
 * MarchingHexahedra / MarchingPyramid / MarchingPrism by cell type
 * (MarchingTetrahedra.cpp:25-286 and the same in the others).  Only their
 * duplicated extraction exists: VertexNormal forces duplication off and a
 * surface without duplication is empty, with the types the mapper leaves.
 * Triangles in cell order and table order; per triangle 3 corners and its
 * normal (v1 - v0) x (v2 - v0).  The colour is kmc_surface_color at isolevel.
 */
kmc_status kmc_extract_unstructured(const kmc_unstructured_volume* volume, double isolevel, int32_t normal_type,
                                    int32_t duplication, int32_t out_memory, void* stream, kmc_surface* out);

/* kvs::SlicePlane over an unstructured volume (SlicePlane.cpp:288-800): as
 * kmc_slice_plane, with node coordinates in place of the grid (no scaling). */
kmc_status kmc_slice_plane_unstructured(const kmc_unstructured_volume* volume, const float coefficients[4],
                                        const uint8_t* color_table, uint32_t color_resolution,
                                        const float* color_range, int32_t out_memory, void* stream,
                                        kmc_surface* out);

/* ------------------------------------------------------------------------- */
/* Volume ingestion: KVSML straight to HBM                                    */
/* ------------------------------------------------------------------------- */

/*
 * kvs::StructuredVolumeObject::read(filename) for KVSML files
 * (StructuredVolumeObject.cpp:171-229, KVSMLStructuredVolumeObject.cpp:159-300,
 * DataArrayTag.cpp:217-420): internal text, external ascii and external binary
 * arrays (either endianness) of char .. double values, uniform / rectilinear /
 * curvilinear grids.  With memory = KMC_MEMORY_DEVICE an external binary array
 * streams from the file into device memory through two pinned staging buffers
 * (disk reads overlap the copies; no host copy of the volume).  *out is a
 * volume ready for kmc_extract: has_min_max_values set (from the file's
 * min_value / max_value, else GetMinMaxValues on the device) and
 * has_min_max_coords set (object_coord, else updateMinMaxCoords of the grid
 * type; non-uniform coordinates themselves are not kept, as the marching-cubes
 * path only reads their extent).  Release with kmc_kvsml_volume_release.
 */
kmc_status kmc_kvsml_read_volume(const char* filename, int32_t memory, void* stream, kmc_volume* out);

/*
 * kvs::StructuredVolumeImporter(filename) for an AVS field file (.fld)
 * (Importer/StructuredVolumeImporter.cpp:37-60, 150-187; FileFormat/AVSField/
 * AVSField.cpp:174-419): the "tag = value" header up to the form feed, the
 * byte / short / integer / float / double values right after the file's first
 * "\f\f", uniform fields only (others: KMC_ERR_FILE, as the importer refuses
 * them).  With memory = KMC_MEMORY_DEVICE the values stream from the file into
 * device memory as kmc_kvsml_read_volume's do.  *out: object coordinates
 * 0 .. dim - 1, min / max values by GetMinMaxValues (computed on the device).
 * Release with kmc_kvsml_volume_release.
 */
kmc_status kmc_avs_read_volume(const char* filename, int32_t memory, void* stream, kmc_volume* out);

/* Release the values of a volume read by kmc_kvsml_read_volume or
 * kmc_avs_read_volume. */
void kmc_kvsml_volume_release(kmc_volume* volume);

/* ------------------------------------------------------------------------- */
/* Output side                                                                 */
/* ------------------------------------------------------------------------- */

/*
 * The surface as a triangle soup, built on the device: 3 corners per triangle
 * gathered through the connections and one normal per triangle (VertexNormal
 * surfaces: (n0 + n1 + n2) / 3.0f of the corners' normals) -- the arrays
 * PolygonExporter<kvs::Stl> builds (Exporter/PolygonExporter.cpp:91-160), and
 * the per-face expansion a renderer of polygon colours / normals needs
 * (Renderer/PolygonRendererGLSL.cpp:105-186).  A duplicated surface is copied.
 * *out: n_vertices = 3 * n_triangles, n_normals = n_triangles, no connections,
 * PolygonNormal, in out_memory (release with kmc_surface_release).
 */
kmc_status kmc_surface_expand(const kmc_surface* surface, int32_t out_memory, void* stream, kmc_surface* out);

/* PolygonExporter<kvs::Stl>(polygon) then Stl::write(filename) with the file
 * type Ascii (ascii != 0) or Binary (FileFormat/STL/Stl.cpp:447-540): the same
 * bytes as the reference. */
kmc_status kmc_export_stl(const kmc_surface* surface, const char* filename, int32_t ascii, void* stream);

/* KVSMLPolygonObject::WritingDataType (FileFormat/KVSML/KVSMLPolygonObject.h:30-35). */
typedef enum kmc_kvsml_writing {
    KMC_KVSML_ASCII = 0,            /* the values inside the XML file */
    KMC_KVSML_EXTERNAL_ASCII = 1,   /* beside it, one ", "-separated text file per array */
    KMC_KVSML_EXTERNAL_BINARY = 2   /* beside it, one raw little-endian file per array */
} kmc_kvsml_writing;

/*
 * PolygonExporter<kvs::KVSMLPolygonObject>(polygon), setWritingDataType(writing)
 * and write(filename) (Exporter/PolygonExporter.cpp:22-70,
 * FileFormat/KVSML/KVSMLPolygonObject.cpp:349-495, KVSML/DataWriter.cpp,
 * DataArrayTag.cpp:74-172, DataValueTag.h:80-105): the same bytes as the
 * reference, in the file and, for the external types, in <base>_coord.dat,
 * _color.dat, _opacity.dat, _normal.dat and _connect.dat in its directory
 * (base: the file name without its last extension).  The polygon's colours
 * are the surface's per-vertex colours when it has them, else `color` (3
 * bytes: the single colour the marching-cubes mappers set; NULL: none);
 * `opacity` is its single opacity (the mappers set 255; -1: none).  A device
 * surface's arrays are copied to the host on `stream` first.
 */
kmc_status kmc_export_kvsml(const kmc_surface* surface, const uint8_t* color, int32_t opacity, const char* filename,
                            int32_t writing, void* stream);

/*
 * PolygonExporter<kvs::Ply>(polygon), setFileType(binary ? Binary : Ascii)
 * and write(filename) (Exporter/PolygonExporter.cpp:178-363,
 * FileFormat/PLY/Ply.cpp:389-489, PlyFile.cpp): ".ply" is appended to a name
 * without it; the header and the vertex records are the reference's bytes --
 * the single colour `color` (when the surface has no per-vertex colours; NULL:
 * none) repeated at every vertex, PolygonNormal normals averaged over each
 * vertex's triangles through the connections (zero without connections), as
 * the exporter does.  Faces follow as the header declares them (a uchar count
 * and int indices); the reference's own writer stops there, writing through a
 * null index pointer (Ply.cpp:475-478).  A device surface's arrays are copied
 * to the host on `stream` first.
 */
kmc_status kmc_export_ply(const kmc_surface* surface, const uint8_t* color, const char* filename, int32_t binary,
                          void* stream);

/* Release the arrays of a surface returned by any kmc_* extraction call.
 * Device arrays go back to the library's device block cache, host arrays
 * (pinned memory) to its pinned block cache (at most 4 GiB kept). */
void kmc_surface_release(kmc_surface* surface);

/*
 * Instrumentation (no reference counterpart): when enabled for the calling
 * thread, kmc_extract / kmc_extract_levels record HIP events on their stream
 * between kernel launches, and kmc_last_timing returns the device time of
 * each kernel of the thread's last such call, summed over its passes.
 */
typedef struct kmc_timing {
    double classify_ms;          /* values -> classification masks */
    double count_ms;             /* per-row vertex / triangle counts (and in-row chunk prefixes) */
    double scan_ms;              /* row prefixes */
    double emit_ms;              /* coordinates, connections, normals: the four below together */
    uint32_t passes;
    double compact_ms;           /* vertex keys and triangle records at their output ranks */
    double vertices_ms;          /* coordinates */
    double triangles_ms;         /* connections and face normals (with duplication: coordinates too) */
    double vertex_normals_ms;    /* VertexNormal sums */
} kmc_timing;

void kmc_set_timing(int32_t enable);
kmc_status kmc_last_timing(kmc_timing* out);

/*
 * GetMinMaxValues<T> (StructuredVolumeObject.cpp:66-111) on the device: the
 * serial scan's result including its NaN behaviour (NaNs are skipped unless
 * the first value is NaN, which then sticks).
 */
kmc_status kmc_volume_min_max(const kmc_volume* volume, void* stream, double* min_value, double* max_value);

/*
 * Surface colour of the reference: TransferFunction::colorMap().at(isolevel)
 * after MapperBase::setRange (MapperBase.cpp:54-83, ColorMap.cpp:1043-1054),
 * for the colour maps named by kmc_colormap.  range_* are the volume's
 * min/max values and type its kmc_type (int8/uint8 use fixed ranges).
 */
typedef enum kmc_colormap {
    KMC_COLORMAP_DEFAULT = 0,              /* ColorMap::DefaultColorMap = BrewerRdBu (ColorMap.cpp:23) */
    KMC_COLORMAP_BREWER_RDBU = 0,
    KMC_COLORMAP_BREWER_SINGLE_HUE_GREYS = 1     /* other maps: KMC_ERR_INVALID_ARGUMENT */
} kmc_colormap;

kmc_status kmc_surface_color(int32_t colormap, uint32_t resolution, int32_t type,
                             double volume_min, double volume_max, double isolevel, uint8_t rgb[3]);

/* The created colour table (3*resolution bytes) of a colour map (ColorMap::create, ColorMap.cpp:913-1019). */
kmc_status kmc_colormap_table(int32_t colormap, uint32_t resolution, uint8_t* rgb_out);

/* ------------------------------------------------------------------------- */
/* Batched form: many volumes of uneven sizes, each with its own isolevels,   */
/* extracted by one hipGraph launch.  No reference counterpart beyond calling */
/* kvs::MarchingCubes once per (volume, isolevel); each surface of a batch is */
/* identical to that call's output.                                           */
/* ------------------------------------------------------------------------- */

/* One job: a volume and the isolevels extracted from it.  Values in device
 * memory are borrowed and must stay valid (and may change) between runs;
 * values in host memory are copied to the device once, at creation. */
typedef struct kmc_job {
    kmc_volume volume;
    const double* isolevels;     /* n_levels values, copied at creation */
    uint32_t n_levels;
    int32_t normal_type;         /* kmc_normal_type */
    int32_t duplication;         /* forced off for KMC_VERTEX_NORMAL, as by the reference */
} kmc_job;

typedef struct kmc_batch kmc_batch;

/* Flags of kmc_batch_create. */
#define KMC_BATCH_AUTOTUNE 1u    /* time the candidate tile shapes on this batch and keep the fastest */

/* What a batch was planned with and what its last run produced. */
typedef struct kmc_batch_info {
    uint32_t n_jobs;             /* jobs as given */
    uint32_t n_surfaces;         /* sum of n_levels */
    uint32_t n_passes;           /* internal passes (a job with more than KMC_MAX_LEVELS_PER_PASS levels takes several) */
    uint32_t n_type_groups;      /* (value type, emit mode) groups: one launch per pass kind and group */
    uint32_t tile_rows;          /* upper bound of node rows per emit tile (autotuned or default) */
    uint32_t queue_capacity;     /* work-queue entries per emit window (autotuned or default) */
    uint32_t graph_nodes;        /* nodes of the captured graph */
    uint32_t recaptures;         /* graph captures so far (output arena growth recaptures) */
    uint64_t nodes_classified;   /* sum over passes of nodes x levels */
    uint64_t n_triangles;        /* sum over surfaces, last run */
    uint64_t arena_bytes;        /* device bytes of the output arena */
    uint64_t work_bytes;         /* device bytes of counts, prefixes and classification masks */
    double last_run_ms;          /* device time of the last run's graph (HIP events) */
    uint32_t flat_emit;          /* 1: flat emit (compaction records, one thread per item); 0: tile teams */
} kmc_batch_info;

/* Plan a batch: validate every job (first failing job decides the status),
 * compute missing min/max values, size the output arena and capture the
 * graph.  stream is used for the planning work; the call returns when done. */
kmc_status kmc_batch_create(const kmc_job* jobs, uint32_t n_jobs, uint32_t flags, void* stream, kmc_batch** out);

/* Run the batch: one graph launch on stream, then wait for it.  The output
 * arena grows (and the graph is captured again) if a run's surfaces do not fit,
 * so results are always complete when the call returns. */
kmc_status kmc_batch_run(kmc_batch* batch, void* stream);

/* Views of the last run's surfaces in device memory, in job order and, within
 * a job, in isolevel order (n_surfaces entries).  The views own nothing
 * (kmc_surface_release only clears them) and stay valid until the next run or
 * kmc_batch_destroy. */
kmc_status kmc_batch_surfaces(const kmc_batch* batch, kmc_surface* outs, uint32_t n_outs);

kmc_status kmc_batch_info_get(const kmc_batch* batch, kmc_batch_info* info);

void kmc_batch_destroy(kmc_batch* batch);

/* One-shot batch: create, run once, copy every surface to storage of its own
 * in out_memory (released with kmc_surface_release), destroy. */
kmc_status kmc_batch_extract(const kmc_job* jobs, uint32_t n_jobs, int32_t out_memory, void* stream,
                             kmc_surface* outs);

/* ------------------------------------------------------------------------- */
/* Serving form: volumes stay resident in HBM; isosurface requests arrive at  */
/* any time and are served by continuous batching.  Worker threads, each with */
/* its own stream, repeatedly take every queued request that fits the free    */
/* slots and the HBM budget and run them as one batch, so requests arriving   */
/* while others are on the GPU join the next batch and batches overlap on the */
/* device.  Each result equals kmc_extract on the same volume and isolevel.   */
/* No reference counterpart (the reference serves one mapper call at a time). */
/* ------------------------------------------------------------------------- */

typedef struct kmc_server kmc_server;

typedef struct kmc_server_config {
    int32_t device;              /* HIP device */
    uint32_t workers;            /* worker threads / streams (batches in flight), default 3 */
    uint32_t max_inflight;       /* requests on the device at once (slots), default 256 */
    uint32_t max_batch;          /* requests per batch, default 64 */
    uint64_t budget_bytes;       /* device bytes for work in flight (masks, counts, outputs), default 16 GiB */
    int32_t out_memory;          /* kmc_memory of returned surfaces */
} kmc_server_config;

typedef struct kmc_server_stats {
    uint64_t submitted, completed, failed;
    uint64_t batches;
    uint64_t triangles;          /* over completed requests */
    uint32_t peak_inflight;      /* most requests admitted at once */
    uint32_t deferred;           /* requests put back because their outputs did not fit the budget */
    uint64_t peak_budget_bytes;  /* most budget bytes held at once */
    double p50_ms, p99_ms, mean_ms, max_ms;  /* latency: submission to result ready */
    double span_s;               /* first submission to last completion */
    double triangles_per_s;      /* triangles / span_s */
    uint64_t peak_batch_device_bytes; /* most device bytes one batch allocated (work arrays, outputs, records) */
} kmc_server_stats;

kmc_status kmc_server_create(const kmc_server_config* config, kmc_server** out);

/* Make a volume resident: host values are copied to the device once; missing
 * min/max values are computed once (they are what every request clamps to and
 * colours by, as the reference caches them in the volume). */
kmc_status kmc_server_register_volume(kmc_server* server, const kmc_volume* volume, uint32_t* volume_id);

/* Queue one request; returns at once with its id.  KMC_ERR_BUDGET if the
 * request alone can never fit the budget. */
kmc_status kmc_server_submit(kmc_server* server, uint32_t volume_id, double isolevel, int32_t normal_type,
                             int32_t duplication, uint64_t* request_id);

/* KMC_ERR_NOT_READY while the request runs; otherwise its status, and on
 * KMC_OK the surface moves to *out (release with kmc_surface_release).  A
 * finished request is forgotten once polled. */
kmc_status kmc_server_poll(kmc_server* server, uint64_t request_id, kmc_surface* out);

/* Like poll, but waits up to timeout_ms (negative: no limit). */
kmc_status kmc_server_wait(kmc_server* server, uint64_t request_id, int64_t timeout_ms, kmc_surface* out);

kmc_status kmc_server_get_stats(kmc_server* server, kmc_server_stats* stats);
kmc_status kmc_server_reset_stats(kmc_server* server);

/* Waits for every queued request, then stops the workers and frees the
 * resident volumes.  Unpolled results are released. */
void kmc_server_destroy(kmc_server* server);

#ifdef __cplusplus
}
#endif

#endif /* KMC_H */
