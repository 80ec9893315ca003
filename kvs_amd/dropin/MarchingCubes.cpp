/*
 * MarchingCubes.cpp -- drop-in implementation of kvs::MarchingCubes on the
 * MI355X marching-cubes library (include/kmc.h, libkmc.so).
 *
 * It replaces Source/Core/Visualization/Mapper/MarchingCubes.cpp of a KVS
 * build and compiles against the unchanged KVS header
 * (Mapper/MarchingCubes.h:22-63): same constructors, same
 * exec(const kvs::ObjectBase*), same outputs in the same kvs::PolygonObject
 * arrays, same error messages.  kvs::Isosurface (Mapper/Isosurface.cpp:176-190)
 * builds a kvs::MarchingCubes for structured volumes and so runs on this
 * implementation unchanged.
 *
 * The surface arrays are adopted without a copy: each kvs::ValueArray holds
 * the library's pinned host buffer through a kvs::SharedPointer whose deleter
 * releases the surface once the last of its arrays goes away
 * (Utility/SharedPointer.h:73-74, Utility/ValueArray.h:121-125).
 */
#include <kvs/MarchingCubes>
#include <kvs/Message>
#include <kvs/Math>
#include <kvs/SharedPointer>
#include <kvs/StructuredVolumeObject>
#include <kvs/ValueArray>

#include <memory>

#include "kmc.h"

namespace
{

/* Owns one surface returned by the library; released with the last array. */
struct SurfaceHolder
{
    kmc_surface surface;
    ~SurfaceHolder() { kmc_surface_release( &surface ); }
};

struct ReleaseSurface
{
    std::shared_ptr<SurfaceHolder> holder;
    template <typename T>
    void operator ()( T* ) { holder.reset(); }
};

template <typename T>
kvs::ValueArray<T> Adopt( const std::shared_ptr<SurfaceHolder>& holder, T* data, size_t size )
{
    return kvs::ValueArray<T>( kvs::SharedPointer<T>( data, ReleaseSurface{ holder } ), size );
}

int TypeOf( const kvs::AnyValueArray& values )
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

/* A device failure: the mapper reports it and exec() returns nullptr. */
thread_local bool t_device_failure = false;

} // namespace

namespace kvs
{

MarchingCubes::MarchingCubes(
    const kvs::StructuredVolumeObject* volume,
    const double isolevel,
    const NormalType normal_type,
    const bool duplication,
    const kvs::TransferFunction& transfer_function ):
    kvs::MapperBase( transfer_function ),
    m_isolevel( isolevel ),
    m_duplication( duplication )
{
    SuperClass::setNormalType( normal_type );
    this->exec( volume );
}

MarchingCubes::SuperClass* MarchingCubes::exec( const kvs::ObjectBase* object )
{
    BaseClass::setSuccess( false );
    if ( !object )
    {
        kvsMessageError( "Input object is NULL." );
        return nullptr;
    }
    const auto* volume = kvs::StructuredVolumeObject::DownCast( object );
    if ( !volume )
    {
        kvsMessageError( "Input object is not structured volume object." );
        return nullptr;
    }
    if ( volume->veclen() != 1 )
    {
        kvsMessageError( "Input volume is not sclar field data." );
        return nullptr;
    }
    // VertexNormal forces shared vertices (reference: MarchingCubes.cpp:69-73).
    if ( SuperClass::normalType() == kvs::PolygonObject::VertexNormal ) { m_duplication = false; }

    t_device_failure = false;
    this->mapping( volume );
    if ( t_device_failure ) { return nullptr; }

    BaseClass::setSuccess( true );
    return this;
}

void MarchingCubes::mapping( const Volume* volume )
{
    BaseClass::attachVolume( volume );

    const int type = TypeOf( volume->values() );
    kmc_volume v{};
    v.values = volume->values().data();
    v.memory = KMC_MEMORY_HOST;
    v.type = type < 0 ? 0 : type;
    v.resolution[0] = volume->resolution().x();
    v.resolution[1] = volume->resolution().y();
    v.resolution[2] = volume->resolution().z();
    v.veclen = 1;
    v.kind = KMC_STRUCTURED;

    // Volume min/max: computed on the device when the volume has none cached,
    // then cached in the volume as updateMinMaxValues would.
    if ( !volume->hasMinMaxValues() && type >= 0 )
    {
        double mn = 0, mx = 0;
        const kmc_status s = kmc_volume_min_max( &v, nullptr, &mn, &mx );
        if ( s != KMC_OK )
        {
            kvsMessageError( "%s", kmc_status_string( s ) );
            t_device_failure = true;
            return;
        }
        volume->setMinMaxValues( mn, mx );
    }
    BaseClass::setRange( volume );
    BaseClass::setMinMaxCoords( volume, this );

    if ( m_duplication )
    {
        SuperClass::setPolygonTypeToTriangle();
        SuperClass::setNormalTypeToPolygon();
        SuperClass::setColorTypeToPolygon();
    }
    else
    {
        SuperClass::setPolygonTypeToTriangle();
        SuperClass::setColorTypeToPolygon();
    }

    const auto min_value = BaseClass::volume()->minValue();
    const auto max_value = BaseClass::volume()->maxValue();
    if ( kvs::Math::Equal( min_value, max_value ) ) { return; }

    if ( type < 0 )
    {
        BaseClass::setSuccess( false );
        kvsMessageError( "Unsupported data type '%s'.", volume->values().typeInfo()->typeName() );
        return;
    }

    v.has_min_max_values = 1;
    v.min_value = min_value;
    v.max_value = max_value;
    v.has_min_max_coords = 1;
    for ( int k = 0; k < 3; ++k )
    {
        v.min_coord[k] = volume->minObjectCoord()[k];
        v.max_coord[k] = volume->maxObjectCoord()[k];
    }

    auto holder = std::make_shared<SurfaceHolder>();
    const kmc_status s = kmc_extract(
        &v, m_isolevel, SuperClass::normalType() == kvs::PolygonObject::VertexNormal ? KMC_VERTEX_NORMAL : KMC_POLYGON_NORMAL,
        m_duplication ? 1 : 0, KMC_MEMORY_HOST, nullptr, &holder->surface );
    if ( s != KMC_OK )
    {
        kvsMessageError( "%s", kmc_status_string( s ) );
        t_device_failure = true;
        return;
    }

    const kmc_surface& out = holder->surface;
    if ( out.n_vertices == 0 ) { return; }   // no coordinates: nothing is set (:236-242, :282-289)

    SuperClass::setCoords( Adopt( holder, out.coords, size_t( 3 * out.n_vertices ) ) );
    if ( !m_duplication )
    {
        SuperClass::setConnections( Adopt( holder, out.connections, size_t( out.n_connections ) ) );
    }
    SuperClass::setColor( BaseClass::transferFunction().colorMap().at( m_isolevel ) );
    SuperClass::setNormals( Adopt( holder, out.normals, size_t( 3 * out.n_normals ) ) );
    SuperClass::setOpacity( 255 );
}

} // end of namespace kvs
