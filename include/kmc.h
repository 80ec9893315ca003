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
    KMC_ERR_NO_DEVICE = 10        /* no HIP device visible */
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

/* Release the arrays of a surface returned by any kmc_* extraction call. */
void kmc_surface_release(kmc_surface* surface);

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
    KMC_COLORMAP_BREWER_SINGLE_HUE_GREYS = 1,
    KMC_COLORMAP_RAINBOW = 2
} kmc_colormap;

kmc_status kmc_surface_color(int32_t colormap, uint32_t resolution, int32_t type,
                             double volume_min, double volume_max, double isolevel, uint8_t rgb[3]);

/* The created colour table (3*resolution bytes) of a colour map (ColorMap::create, ColorMap.cpp:913-1019). */
kmc_status kmc_colormap_table(int32_t colormap, uint32_t resolution, uint8_t* rgb_out);

#ifdef __cplusplus
}
#endif

#endif /* KMC_H */
