/*
 * MarchingTetrahedra.cpp -- drop-in implementation of kvs::MarchingTetrahedra on the MI355X
 * library (include/kmc.h: kmc_extract_unstructured).
 *
 * It replaces Source/Core/Visualization/Mapper/MarchingTetrahedra.cpp of a KVS build and
 * compiles against the unchanged header (MarchingTetrahedra.h): same constructor, same
 * exec(const kvs::ObjectBase*), the same messages, types and arrays
 * (MarchingTetrahedra.cpp:25-136).  kvs::Isosurface builds a kvs::MarchingTetrahedra for
 * tetrahedral volumes (Isosurface.cpp:197-211) and so runs on this
 * implementation unchanged.  The volume's min/max values and coordinates are
 * computed on the device when it has none cached, then cached in it as the
 * reference's update functions would.
 */
#include <kvs/MarchingTetrahedra>
#include <kvs/Message>
#include <kvs/Math>
#include <kvs/UnstructuredVolumeObject>

#include "DropinSurface.h"

namespace
{
/* A device failure: the mapper reports it and exec() returns nullptr. */
thread_local bool t_device_failure = false;
} // namespace

namespace kvs
{

MarchingTetrahedra::MarchingTetrahedra(
    const kvs::UnstructuredVolumeObject* volume,
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

MarchingTetrahedra::SuperClass* MarchingTetrahedra::exec( const kvs::ObjectBase* object )
{
    if ( !object )
    {
        BaseClass::setSuccess( false );
        kvsMessageError("Input object is NULL.");
        return nullptr;
    }
    const auto* volume = kvs::UnstructuredVolumeObject::DownCast( object );
    if ( !volume )
    {
        BaseClass::setSuccess( false );
        kvsMessageError("Input object is not unstructured volume object.");
        return nullptr;
    }
    if ( volume->veclen() != 1 )
    {
        BaseClass::setSuccess( false );
        kvsMessageError("Input volume is not sclar field data.");
        return nullptr;
    }
    if ( volume->cellType() != kvs::UnstructuredVolumeObject::Tetrahedra )
    {
        BaseClass::setSuccess( false );
        kvsMessageError("Input volume is not tetrahedra-cell data.");
        return nullptr;
    }
    // VertexNormal forces shared vertices (MarchingTetrahedra.cpp:77-81); the
    // reference has no shared-vertex extraction, so the surface stays empty.
    if ( SuperClass::normalType() == kvs::PolygonObject::VertexNormal ) { m_duplication = false; }

    t_device_failure = false;
    this->mapping( volume );
    if ( t_device_failure ) { return nullptr; }
    return this;
}

void MarchingTetrahedra::mapping( const kvs::UnstructuredVolumeObject* volume )
{
    BaseClass::attachVolume( volume );

    // The extraction first: it also yields the volume's min/max (values and
    // coordinates, in connection order), cached before setRange reads them.
    const int type = kmc_dropin::TypeOf( volume->values() );
    auto holder = std::make_shared<kmc_dropin::SurfaceHolder>();
    if ( type >= 0 && volume->numberOfNodes() > 0 )
    {
        kmc_unstructured_volume v = kmc_dropin::UnstructuredView( volume, type );
        const int normal = SuperClass::normalType() == kvs::PolygonObject::VertexNormal ? KMC_VERTEX_NORMAL : KMC_POLYGON_NORMAL;
        const kmc_status s = kmc_extract_unstructured(
            &v, m_isolevel, normal, m_duplication ? 1 : 0, KMC_MEMORY_HOST, nullptr, &holder->surface );
        if ( s != KMC_OK )
        {
            kvsMessageError( "%s", kmc_status_string( s ) );
            t_device_failure = true;
            return;
        }
        kmc_dropin::CacheMinMax( volume, holder->surface );
    }
    BaseClass::setRange( volume );
    BaseClass::setMinMaxCoords( volume, this );

    if ( m_duplication )
    {
        SuperClass::setPolygonType( kvs::PolygonObject::Triangle );
        SuperClass::setColorType( kvs::PolygonObject::PolygonColor );
        SuperClass::setNormalType( kvs::PolygonObject::PolygonNormal );
    }

    const auto min_value = BaseClass::volume()->minValue();
    const auto max_value = BaseClass::volume()->maxValue();
    if ( kvs::Math::Equal( min_value, max_value ) ) { return; }

    if ( type < 0 )
    {
        BaseClass::setSuccess( false );
        kvsMessageError("Unsupported data type '%s'.", volume->values().typeInfo()->typeName() );
        return;
    }

    const kmc_surface& out = holder->surface;
    if ( out.n_vertices > 0 )
    {
        SuperClass::setCoords( kmc_dropin::Adopt( holder, out.coords, size_t( 3 * out.n_vertices ) ) );
        SuperClass::setColor( BaseClass::transferFunction().colorMap().at( m_isolevel ) );
        SuperClass::setNormals( kmc_dropin::Adopt( holder, out.normals, size_t( 3 * out.n_normals ) ) );
        SuperClass::setOpacity( 255 );
    }
}

} // end of namespace kvs
