/*
 * DropinSurface.h -- what the drop-in mapper classes (kvs_amd/dropin/*.cpp)
 * share: adopting a library surface's arrays into kvs::ValueArrays without a
 * copy, the value-type map, and the unstructured extraction with the
 * volume's min/max cached from the device as the reference's
 * updateMinMaxValues / updateMinMaxCoords would cache them
 * (UnstructuredVolumeObject.cpp:123-141, 395-467).
 */
#pragma once
#include <kvs/AnyValueArray>
#include <kvs/SharedPointer>
#include <kvs/UnstructuredVolumeObject>
#include <kvs/ValueArray>
#include <kvs/Vector3>

#include <memory>

#include "kmc.h"

namespace kmc_dropin
{

/* Owns one surface returned by the library; released with the last array. */
struct SurfaceHolder
{
    kmc_surface surface{};
    ~SurfaceHolder() { kmc_surface_release( &surface ); }
};

struct ReleaseSurface
{
    std::shared_ptr<SurfaceHolder> holder;
    template <typename T>
    void operator ()( T* ) { holder.reset(); }
};

/* A kvs::ValueArray over the library's pinned buffer (Utility/SharedPointer.h:73-74,
 * Utility/ValueArray.h:121-125): the surface is released with the last array. */
template <typename T>
inline kvs::ValueArray<T> Adopt( const std::shared_ptr<SurfaceHolder>& holder, T* data, size_t size )
{
    return kvs::ValueArray<T>( kvs::SharedPointer<T>( data, ReleaseSurface{ holder } ), size );
}

/* kvs::Type::TypeID -> kmc_type (-1: not a supported value type). */
inline int TypeOf( const kvs::AnyValueArray& values )
{
    switch ( values.typeID() )
    {
    case kvs::Type::TypeInt8: return KMC_INT8;
    case kvs::Type::TypeInt16: return KMC_INT16;
    case kvs::Type::TypeInt32: return KMC_INT32;
    case kvs::Type::TypeInt64: return KMC_INT64;
    case kvs::Type::TypeUInt8: return KMC_UINT8;
    case kvs::Type::TypeUInt16: return KMC_UINT16;
    case kvs::Type::TypeUInt32: return KMC_UINT32;
    case kvs::Type::TypeUInt64: return KMC_UINT64;
    case kvs::Type::TypeReal32: return KMC_REAL32;
    case kvs::Type::TypeReal64: return KMC_REAL64;
    default: return -1;
    }
}

/* The library's view of an unstructured volume (host arrays), with the
 * volume's cached min/max values and object coordinates when it has them. */
inline kmc_unstructured_volume UnstructuredView( const kvs::UnstructuredVolumeObject* volume, int type )
{
    kmc_unstructured_volume v{};
    v.values = volume->values().data();
    v.coords = volume->coords().data();
    v.connections = volume->connections().data();
    v.memory = KMC_MEMORY_HOST;
    v.type = type;
    v.n_nodes = volume->numberOfNodes();
    v.n_cells = volume->numberOfCells();
    v.cell_type = int32_t( volume->cellType() );
    v.veclen = 1;
    v.has_min_max_values = volume->hasMinMaxValues() ? 1 : 0;
    v.min_value = volume->hasMinMaxValues() ? volume->minValue() : 0.0;
    v.max_value = volume->hasMinMaxValues() ? volume->maxValue() : 0.0;
    v.has_min_max_coords = volume->hasMinMaxObjectCoords() ? 1 : 0;
    for ( int k = 0; k < 3; ++k )
    {
        v.min_coord[k] = volume->minObjectCoord()[k];
        v.max_coord[k] = volume->maxObjectCoord()[k];
    }
    return v;
}

/* Caches in the volume the min/max values and object coordinates the library
 * computed for surface s (in connection order, as the reference's
 * updateMinMaxValues / updateMinMaxCoords compute them), so that
 * MapperBase::setRange / setMinMaxCoords find them and scan nothing. */
inline void CacheMinMax( const kvs::UnstructuredVolumeObject* volume, const kmc_surface& s )
{
    if ( !volume->hasMinMaxValues() ) { volume->setMinMaxValues( s.min_value, s.max_value ); }
    if ( !volume->hasMinMaxObjectCoords() && volume->numberOfCells() > 0 )
    {
        // WARNING: removes constness, as MapperBase::setMinMaxCoords does (MapperBase.cpp:103-108).
        auto* v = const_cast<kvs::UnstructuredVolumeObject*>( volume );
        const kvs::Vec3 mn( s.min_coord[0], s.min_coord[1], s.min_coord[2] );
        const kvs::Vec3 mx( s.max_coord[0], s.max_coord[1], s.max_coord[2] );
        v->setMinMaxObjectCoords( mn, mx );
        if ( !v->hasMinMaxExternalCoords() ) { v->setMinMaxExternalCoords( mn, mx ); }
    }
}

} // namespace kmc_dropin
