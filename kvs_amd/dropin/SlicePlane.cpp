/*
 * SlicePlane.cpp -- drop-in implementation of kvs::SlicePlane on the MI355X
 * library (include/kmc.h: kmc_slice_plane, kmc_slice_plane_unstructured).
 *
 * It replaces Source/Core/Visualization/Mapper/SlicePlane.cpp of a KVS build
 * and compiles against the unchanged header (SlicePlane.h): same
 * constructors, same exec(const kvs::ObjectBase*), the same messages, types
 * and arrays (SlicePlane.cpp:26-150, 277-283 and the unstructured variants'
 * ends).  The triangles, their corner colours (the transfer function's own
 * colour map table and range) and normals are computed on the device; the
 * arrays are adopted without a copy.  Volume min/max values and coordinates
 * missing from the volume are computed on the device and cached in it first.
 */
#include <kvs/SlicePlane>
#include <kvs/Message>
#include <kvs/StructuredVolumeObject>
#include <kvs/UnstructuredVolumeObject>

#include "DropinSurface.h"

namespace
{
/* A device failure: the mapper reports it and exec() returns nullptr. */
thread_local bool t_device_failure = false;
} // namespace

namespace kvs
{

SlicePlane::SlicePlane(
    const kvs::VolumeObjectBase* volume,
    const kvs::Vec4& coefficients,
    const kvs::TransferFunction& transfer_function ):
    kvs::MapperBase( transfer_function ),
    kvs::PolygonObject()
{
    this->setPlane( coefficients );
    this->exec( volume );
}

SlicePlane::SlicePlane(
    const kvs::VolumeObjectBase* volume,
    const kvs::Vec3& point,
    const kvs::Vec3& normal,
    const kvs::TransferFunction& transfer_function ):
    kvs::MapperBase( transfer_function ),
    kvs::PolygonObject()
{
    this->setPlane( point, normal );
    this->exec( volume );
}

SlicePlane::SuperClass* SlicePlane::exec( const kvs::ObjectBase* object )
{
    if ( !object )
    {
        BaseClass::setSuccess( false );
        kvsMessageError("Input object is NULL.");
        return NULL;
    }
    const auto* volume = kvs::VolumeObjectBase::DownCast( object );
    if ( !volume )
    {
        BaseClass::setSuccess( false );
        kvsMessageError("Input object is not volume dat.");
        return NULL;
    }
    t_device_failure = false;
    this->mapping( volume );
    if ( t_device_failure ) { return NULL; }
    return this;
}

void SlicePlane::mapping( const kvs::VolumeObjectBase* volume )
{
    if ( volume->veclen() != 1 )
    {
        BaseClass::setSuccess( false );
        kvsMessageError("Input volume is not a sclar field data.");
        return;
    }
    BaseClass::attachVolume( volume );

    const int type = kmc_dropin::TypeOf( volume->values() );
    const float coef[4] = { m_coef[0], m_coef[1], m_coef[2], m_coef[3] };
    auto holder = std::make_shared<kmc_dropin::SurfaceHolder>();
    auto fail = [&] ( kmc_status s )
    {
        kvsMessageError( "%s", kmc_status_string( s ) );
        t_device_failure = true;
    };

    bool extracted = false;
    if ( volume->volumeType() == kvs::VolumeObjectBase::Structured )
    {
        const auto* sv = kvs::StructuredVolumeObject::DownCast( volume );
        kmc_volume v{};
        v.values = sv->values().data();
        v.memory = KMC_MEMORY_HOST;
        v.type = type < 0 ? 0 : type;
        v.resolution[0] = sv->resolution().x();
        v.resolution[1] = sv->resolution().y();
        v.resolution[2] = sv->resolution().z();
        v.veclen = 1;
        v.kind = KMC_STRUCTURED;
        if ( !sv->hasMinMaxValues() && type >= 0 && sv->numberOfNodes() > 0 )
        {
            double mn = 0, mx = 0;
            const kmc_status s = kmc_volume_min_max( &v, nullptr, &mn, &mx );
            if ( s != KMC_OK ) { fail( s ); return; }
            sv->setMinMaxValues( mn, mx );
        }
        BaseClass::setRange( volume );
        BaseClass::setMinMaxCoords( volume, this );
        if ( type >= 0 )
        {
            const auto& cmap = BaseClass::transferFunction().colorMap();
            const float range[2] = { cmap.minValue(), cmap.maxValue() };
            v.has_min_max_values = 1;
            v.min_value = sv->minValue();
            v.max_value = sv->maxValue();
            v.has_min_max_coords = 1;
            for ( int k = 0; k < 3; ++k )
            {
                v.min_coord[k] = sv->minObjectCoord()[k];
                v.max_coord[k] = sv->maxObjectCoord()[k];
            }
            const kmc_status s = kmc_slice_plane( &v, coef, cmap.table().data(), uint32_t( cmap.resolution() ), range,
                                                  KMC_MEMORY_HOST, nullptr, &holder->surface );
            if ( s != KMC_OK ) { fail( s ); return; }
            extracted = true;
        }
    }
    else
    {
        const auto* uv = kvs::UnstructuredVolumeObject::DownCast( volume );
        const int cell = int( uv->cellType() );
        const bool has_table = cell == KMC_CELL_TETRAHEDRA || cell == KMC_CELL_HEXAHEDRA ||
                               cell == KMC_CELL_PYRAMID || cell == KMC_CELL_PRISM;
        if ( type >= 0 && has_table && uv->numberOfNodes() > 0 )
        {
            // the range MapperBase::setRange will give (NULL: from the volume's
            // min/max, which the call computes when the volume has none)
            const auto& tfunc = BaseClass::transferFunction();
            const auto& cmap = tfunc.colorMap();
            const float range[2] = { cmap.minValue(), cmap.maxValue() };
            kmc_unstructured_volume v = kmc_dropin::UnstructuredView( uv, type );
            const kmc_status s = kmc_slice_plane_unstructured(
                &v, coef, cmap.table().data(), uint32_t( cmap.resolution() ), tfunc.hasRange() ? range : nullptr,
                KMC_MEMORY_HOST, nullptr, &holder->surface );
            if ( s != KMC_OK ) { fail( s ); return; }
            kmc_dropin::CacheMinMax( uv, holder->surface );
            extracted = true;
        }
        BaseClass::setRange( volume );
        BaseClass::setMinMaxCoords( volume, this );
        if ( !has_table && type >= 0 ) { return; }   // no marching table for the cell type: nothing is set
    }

    if ( type < 0 )
    {
        BaseClass::setSuccess( false );
        kvsMessageError("Unsupported data type '%s'.", volume->values().typeInfo()->typeName() );
        return;
    }

    // the arrays (empty when nothing was cut) and the types, as every
    // extract_*plane leaves them
    const kmc_surface& out = holder->surface;
    if ( extracted && out.n_vertices > 0 )
    {
        SuperClass::setCoords( kmc_dropin::Adopt( holder, out.coords, size_t( 3 * out.n_vertices ) ) );
        SuperClass::setColors( kmc_dropin::Adopt( holder, out.colors, size_t( 3 * out.n_colors ) ) );
        SuperClass::setNormals( kmc_dropin::Adopt( holder, out.normals, size_t( 3 * out.n_normals ) ) );
    }
    else
    {
        SuperClass::setCoords( kvs::ValueArray<kvs::Real32>() );
        SuperClass::setColors( kvs::ValueArray<kvs::UInt8>() );
        SuperClass::setNormals( kvs::ValueArray<kvs::Real32>() );
    }
    SuperClass::setOpacity( 255 );
    SuperClass::setPolygonType( kvs::PolygonObject::Triangle );
    SuperClass::setColorType( kvs::PolygonObject::VertexColor );
    SuperClass::setNormalType( kvs::PolygonObject::PolygonNormal );
}

} // end of namespace kvs
