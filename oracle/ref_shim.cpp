/*
 * ref_shim.cpp -- test infrastructure only.
 *
 * A C ABI over the reference's own kvs::MarchingCubes / kvs::Isosurface,
 * compiled together with the reference's sources (ref_sources.txt) into
 * oracle/_ref/libkvsref.so by oracle/Makefile.  It lets the tests check the
 * oracle restatement (mc_oracle.c) and the colour restatement against the
 * real reference, and lets bench.py time the real reference as the CPU
 * baseline.  Nothing here is part of the product.
 */
#include <kvs/MarchingCubes>
#include <kvs/MarchingCubesTable>
#include <kvs/Isosurface>
#include <kvs/StructuredVolumeObject>
#include <kvs/TransferFunction>
#include <kvs/AVSField>
#include <kvs/StructuredVolumeImporter>
#include <kvs/ColorMap>
#include <kvs/ValueArray>
#include <kvs/AnyValueArray>
#include <kvs/Type>
#include <kvs/HydrogenVolumeData>
#include <kvs/SlicePlane>
#include <kvs/PolygonExporter>
#include <kvs/Stl>
#include <kvs/KVSMLPolygonObject>
#include <kvs/Ply>
#include <kvs/UnstructuredVolumeObject>
#include <kvs/MarchingTetrahedraTable>
#include <kvs/MarchingHexahedraTable>
#include <kvs/MarchingPyramidTable>
#include <kvs/MarchingPrismTable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <chrono>

namespace
{

template <typename T>
kvs::AnyValueArray MakeValues( const void* data, size_t n )
{
    kvs::ValueArray<T> a( n );
    std::memcpy( a.data(), data, n * sizeof( T ) );
    return kvs::AnyValueArray( a );
}

kvs::AnyValueArray MakeAny( const void* data, int dtype, size_t n )
{
    switch ( dtype )
    {
    case 0: return MakeValues<kvs::Int8>( data, n );
    case 1: return MakeValues<kvs::Int16>( data, n );
    case 2: return MakeValues<kvs::Int32>( data, n );
    case 3: return MakeValues<kvs::Int64>( data, n );
    case 4: return MakeValues<kvs::UInt8>( data, n );
    case 5: return MakeValues<kvs::UInt16>( data, n );
    case 6: return MakeValues<kvs::UInt32>( data, n );
    case 7: return MakeValues<kvs::UInt64>( data, n );
    case 8: return MakeValues<kvs::Real32>( data, n );
    default: return MakeValues<kvs::Real64>( data, n );
    }
}

kvs::TransferFunction MakeTF( int tf_kind, int resolution )
{
    // 0: TransferFunction(resolution) (the reference default, BrewerRdBu);
    // 1: BrewerSingleHueGreys(resolution); 2: Rainbow(resolution).
    if ( tf_kind == 1 ) return kvs::TransferFunction( kvs::ColorMap::BrewerSingleHueGreys( resolution ) );
    if ( tf_kind == 2 ) return kvs::TransferFunction( kvs::ColorMap::Rainbow( resolution ) );
    return kvs::TransferFunction( resolution );
}

template <typename A>
void* CopyOut( const A& a, uint64_t* n )
{
    *n = a.size();
    if ( a.size() == 0 ) return nullptr;
    void* p = std::malloc( a.byteSize() );
    std::memcpy( p, a.data(), a.byteSize() );
    return p;
}

} // namespace

extern "C" {

struct ref_volume
{
    const void* values;
    int32_t dtype;
    uint32_t dims[3];
    int32_t has_min_max_values;
    double min_value, max_value;
    int32_t has_min_max_coords;
    float min_coord[3], max_coord[3];
};

struct ref_mesh
{
    float* coords; uint64_t n_coords;
    uint32_t* connections; uint64_t n_connections;
    float* normals; uint64_t n_normals;
    uint8_t* colors; uint64_t n_colors;
    uint8_t* opacities; uint64_t n_opacities;
    int32_t polygon_type, normal_type, color_type;
    int32_t success;
    double volume_min_value, volume_max_value;   // after the call (cached by the mapper)
    float min_obj[3], max_obj[3];                // polygon's min/max object coords
    double seconds;                              // wall time of the mapper constructor alone
};

static kvs::StructuredVolumeObject* MakeVolume( const ref_volume* v )
{
    auto* vol = new kvs::StructuredVolumeObject();
    vol->setGridTypeToUniform();
    vol->setVeclen( 1 );
    vol->setResolution( kvs::Vec3u( v->dims[0], v->dims[1], v->dims[2] ) );
    const size_t n = size_t( v->dims[0] ) * v->dims[1] * v->dims[2];
    vol->setValues( MakeAny( v->values, v->dtype, n ) );
    if ( v->has_min_max_values ) vol->setMinMaxValues( v->min_value, v->max_value );
    if ( v->has_min_max_coords )
    {
        vol->setMinMaxObjectCoords(
            kvs::Vec3( v->min_coord[0], v->min_coord[1], v->min_coord[2] ),
            kvs::Vec3( v->max_coord[0], v->max_coord[1], v->max_coord[2] ) );
    }
    return vol;
}

static void Harvest( const kvs::PolygonObject* p, const kvs::StructuredVolumeObject* vol, int success, ref_mesh* out )
{
    std::memset( out, 0, sizeof( *out ) );
    out->coords = static_cast<float*>( CopyOut( p->coords(), &out->n_coords ) );
    out->connections = static_cast<uint32_t*>( CopyOut( p->connections(), &out->n_connections ) );
    out->normals = static_cast<float*>( CopyOut( p->normals(), &out->n_normals ) );
    out->colors = static_cast<uint8_t*>( CopyOut( p->colors(), &out->n_colors ) );
    out->opacities = static_cast<uint8_t*>( CopyOut( p->opacities(), &out->n_opacities ) );
    out->polygon_type = p->polygonType();
    out->normal_type = p->normalType();
    out->color_type = p->colorType();
    out->success = success;
    out->volume_min_value = vol->minValue();
    out->volume_max_value = vol->maxValue();
    for ( int k = 0; k < 3; ++k )
    {
        out->min_obj[k] = p->minObjectCoord()[k];
        out->max_obj[k] = p->maxObjectCoord()[k];
    }
}

// kvs::MarchingCubes( volume, isolevel, normal_type, duplication, tf ).
int ref_marching_cubes( const ref_volume* v, double isolevel, int normal_type, int duplication,
                        int tf_kind, int tf_resolution, ref_mesh* out )
{
    auto* vol = MakeVolume( v );
    const auto tf = MakeTF( tf_kind, tf_resolution );
    const auto t0 = std::chrono::steady_clock::now();
    auto* mc = new kvs::MarchingCubes(
        vol, isolevel, kvs::PolygonObject::NormalType( normal_type ), duplication != 0, tf );
    const auto t1 = std::chrono::steady_clock::now();
    Harvest( mc, vol, mc->isSuccess() ? 1 : 0, out );
    out->seconds = std::chrono::duration<double>( t1 - t0 ).count();
    delete mc;
    delete vol;
    return 0;
}

// kvs::Isosurface( volume, isolevel, normal_type, duplication, tf ) when
// use_default_level == 0, else kvs::Isosurface( volume, normal_type ).
int ref_isosurface( const ref_volume* v, int use_default_level, double isolevel, int normal_type,
                    int duplication, int tf_kind, int tf_resolution, ref_mesh* out )
{
    auto* vol = MakeVolume( v );
    kvs::Isosurface* s = nullptr;
    if ( use_default_level )
        s = new kvs::Isosurface( vol, kvs::PolygonObject::NormalType( normal_type ) );
    else
        s = new kvs::Isosurface( vol, isolevel, kvs::PolygonObject::NormalType( normal_type ),
                                 duplication != 0, MakeTF( tf_kind, tf_resolution ) );
    Harvest( s, vol, s->isSuccess() ? 1 : 0, out );
    delete s;
    delete vol;
    return 0;
}

void ref_mesh_free( ref_mesh* m )
{
    std::free( m->coords ); std::free( m->connections ); std::free( m->normals );
    std::free( m->colors ); std::free( m->opacities );
    std::memset( m, 0, sizeof( *m ) );
}

// GetMinMaxValues<T> through StructuredVolumeObject::updateMinMaxValues.
int ref_min_max( const void* values, int dtype, const uint32_t dims[3], double* mn, double* mx )
{
    ref_volume v{};
    v.values = values; v.dtype = dtype;
    v.dims[0] = dims[0]; v.dims[1] = dims[1]; v.dims[2] = dims[2];
    auto* vol = MakeVolume( &v );
    vol->updateMinMaxValues();
    *mn = vol->minValue(); *mx = vol->maxValue();
    delete vol;
    return 0;
}

// ColorMap::at( value ) of the given transfer function after
// TransferFunction::setRange( range_min, range_max ).
int ref_tf_color( int tf_kind, int tf_resolution, float range_min, float range_max, float value, uint8_t rgb[3] )
{
    auto tf = MakeTF( tf_kind, tf_resolution );
    tf.setRange( range_min, range_max );
    const kvs::RGBColor c = tf.colorMap().at( value );
    rgb[0] = c.r(); rgb[1] = c.g(); rgb[2] = c.b();
    return 0;
}

// The created colour table of a transfer function (3*resolution bytes).
int ref_tf_table( int tf_kind, int tf_resolution, uint8_t* out )
{
    auto tf = MakeTF( tf_kind, tf_resolution );
    const auto& t = tf.colorMap().table();
    std::memcpy( out, t.data(), t.size() );
    return int( t.size() );
}

// kvs::HydrogenVolumeData( dims ): the reference's own uint8 test volume
// (Visualization/Data/HydrogenVolumeData.cpp:51-98).
int ref_hydrogen( const uint32_t dims[3], uint8_t* out )
{
    kvs::HydrogenVolumeData h( kvs::Vec3u( dims[0], dims[1], dims[2] ) );
    const auto& v = h.values();
    std::memcpy( out, v.data(), v.byteSize() );
    return int( v.byteSize() );
}

// kvs::SlicePlane( volume, coefficients, tf ) over a structured volume.
int ref_slice_plane( const ref_volume* v, const float coef[4], int tf_kind, int tf_resolution, ref_mesh* out )
{
    auto* vol = MakeVolume( v );
    const auto t0 = std::chrono::steady_clock::now();
    auto* sp = new kvs::SlicePlane( vol, kvs::Vec4( coef[0], coef[1], coef[2], coef[3] ), MakeTF( tf_kind, tf_resolution ) );
    const auto t1 = std::chrono::steady_clock::now();
    Harvest( sp, vol, sp->isSuccess() ? 1 : 0, out );
    out->seconds = std::chrono::duration<double>( t1 - t0 ).count();
    delete sp;
    delete vol;
    return 0;
}

// kvs::SlicePlane( volume, point, normal, tf ): the coefficients the
// point-normal form stores (SlicePlane.h:51: { normal, -point.dot( normal ) }).
void ref_slice_plane_coefficients( const float point[3], const float normal[3], float coef[4] )
{
    kvs::SlicePlane sp;
    sp.setPlane( kvs::Vec3( point[0], point[1], point[2] ), kvs::Vec3( normal[0], normal[1], normal[2] ) );
    const kvs::Vec3 p( point[0], point[1], point[2] ), n( normal[0], normal[1], normal[2] );
    coef[0] = n[0]; coef[1] = n[1]; coef[2] = n[2]; coef[3] = -p.dot( n );
}

struct ref_uvolume
{
    const void* values;
    int32_t dtype;
    uint64_t nnodes, ncells;
    int32_t cell_type;                 // kvs::UnstructuredVolumeObject::CellType
    const float* coords;               // 3 * nnodes
    const uint32_t* connections;       // nodes per cell * ncells
    int32_t has_min_max_values;
    double min_value, max_value;
};

static kvs::UnstructuredVolumeObject* MakeUVolume( const ref_uvolume* v )
{
    auto* vol = new kvs::UnstructuredVolumeObject();
    vol->setCellType( kvs::UnstructuredVolumeObject::CellType( v->cell_type ) );
    vol->setVeclen( 1 );
    vol->setNumberOfNodes( v->nnodes );
    vol->setNumberOfCells( v->ncells );
    vol->setValues( MakeAny( v->values, v->dtype, v->nnodes ) );
    kvs::ValueArray<kvs::Real32> coords( 3 * v->nnodes );
    std::memcpy( coords.data(), v->coords, 3 * v->nnodes * sizeof( float ) );
    vol->setCoords( coords );
    const size_t per = vol->numberOfCellNodes();
    kvs::ValueArray<kvs::UInt32> conn( per * v->ncells );
    std::memcpy( conn.data(), v->connections, per * v->ncells * sizeof( uint32_t ) );
    vol->setConnections( conn );
    if ( v->has_min_max_values ) vol->setMinMaxValues( v->min_value, v->max_value );
    vol->updateMinMaxCoords();
    return vol;
}

static void HarvestU( const kvs::PolygonObject* p, const kvs::UnstructuredVolumeObject* vol, int success, ref_mesh* out )
{
    std::memset( out, 0, sizeof( *out ) );
    out->coords = static_cast<float*>( CopyOut( p->coords(), &out->n_coords ) );
    out->connections = static_cast<uint32_t*>( CopyOut( p->connections(), &out->n_connections ) );
    out->normals = static_cast<float*>( CopyOut( p->normals(), &out->n_normals ) );
    out->colors = static_cast<uint8_t*>( CopyOut( p->colors(), &out->n_colors ) );
    out->opacities = static_cast<uint8_t*>( CopyOut( p->opacities(), &out->n_opacities ) );
    out->polygon_type = p->polygonType();
    out->normal_type = p->normalType();
    out->color_type = p->colorType();
    out->success = success;
    out->volume_min_value = vol->minValue();
    out->volume_max_value = vol->maxValue();
    for ( int k = 0; k < 3; ++k )
    {
        out->min_obj[k] = p->minObjectCoord()[k];
        out->max_obj[k] = p->maxObjectCoord()[k];
    }
}

// kvs::Isosurface( volume, isolevel, normal_type, duplication, tf ) over an
// unstructured volume: MarchingTetrahedra / Hexahedra / Pyramid / Prism by
// cell type (Isosurface.cpp:192-257).
int ref_isosurface_unstructured( const ref_uvolume* v, double isolevel, int normal_type, int duplication,
                                 int tf_kind, int tf_resolution, ref_mesh* out )
{
    auto* vol = MakeUVolume( v );
    const auto t0 = std::chrono::steady_clock::now();
    auto* s = new kvs::Isosurface( vol, isolevel, kvs::PolygonObject::NormalType( normal_type ),
                                   duplication != 0, MakeTF( tf_kind, tf_resolution ) );
    const auto t1 = std::chrono::steady_clock::now();
    HarvestU( s, vol, s->isSuccess() ? 1 : 0, out );
    out->seconds = std::chrono::duration<double>( t1 - t0 ).count();
    delete s;
    delete vol;
    return 0;
}

// kvs::SlicePlane( volume, coefficients, tf ) over an unstructured volume.
int ref_slice_plane_unstructured( const ref_uvolume* v, const float coef[4], int tf_kind, int tf_resolution,
                                  ref_mesh* out )
{
    auto* vol = MakeUVolume( v );
    auto* sp = new kvs::SlicePlane( vol, kvs::Vec4( coef[0], coef[1], coef[2], coef[3] ), MakeTF( tf_kind, tf_resolution ) );
    HarvestU( sp, vol, sp->isSuccess() ? 1 : 0, out );
    delete sp;
    delete vol;
    return 0;
}

// The unstructured marching tables (triangle edge lists and edge endpoints)
// for pinning transcriptions: cell 0 tetrahedra [16][7] / [6][2], 1
// hexahedra [256][16] / [12][2], 2 pyramid [36][13] / [8][2], 3 prism
// [64][10] / [9][2].  Returns the number of int32 written to tri.
int ref_unstructured_tables( int cell, int32_t* tri, int32_t* vert )
{
    switch ( cell )
    {
    case 0:
        std::memcpy( tri, kvs::MarchingTetrahedraTable::TriangleID, sizeof( kvs::MarchingTetrahedraTable::TriangleID ) );
        std::memcpy( vert, kvs::MarchingTetrahedraTable::VertexID, sizeof( kvs::MarchingTetrahedraTable::VertexID ) );
        return int( sizeof( kvs::MarchingTetrahedraTable::TriangleID ) / 4 );
    case 1:
        std::memcpy( tri, kvs::MarchingHexahedraTable::TriangleID, sizeof( kvs::MarchingHexahedraTable::TriangleID ) );
        std::memcpy( vert, kvs::MarchingHexahedraTable::VertexID, sizeof( kvs::MarchingHexahedraTable::VertexID ) );
        return int( sizeof( kvs::MarchingHexahedraTable::TriangleID ) / 4 );
    case 2:
        std::memcpy( tri, kvs::MarchingPyramidTable::TriangleID, sizeof( kvs::MarchingPyramidTable::TriangleID ) );
        std::memcpy( vert, kvs::MarchingPyramidTable::VertexID, sizeof( kvs::MarchingPyramidTable::VertexID ) );
        return int( sizeof( kvs::MarchingPyramidTable::TriangleID ) / 4 );
    default:
        std::memcpy( tri, kvs::MarchingPrismTable::TriangleID, sizeof( kvs::MarchingPrismTable::TriangleID ) );
        std::memcpy( vert, kvs::MarchingPrismTable::VertexID, sizeof( kvs::MarchingPrismTable::VertexID ) );
        return int( sizeof( kvs::MarchingPrismTable::TriangleID ) / 4 );
    }
}

// StructuredVolumeObject::write( filename, ascii, external ) (KVSML) of a
// uniform volume, or of a rectilinear one when rect_coords is given
// (nx + ny + nz floats).
int ref_write_kvsml_volume( const ref_volume* v, const float* rect_coords, const char* filename, int ascii,
                            int external )
{
    auto* vol = MakeVolume( v );
    if ( rect_coords )
    {
        vol->setGridTypeToRectilinear();
        const size_t n = size_t( v->dims[0] ) + v->dims[1] + v->dims[2];
        kvs::ValueArray<kvs::Real32> c( n );
        std::memcpy( c.data(), rect_coords, n * sizeof( float ) );
        vol->setCoords( c );
    }
    if ( !v->has_min_max_values ) vol->updateMinMaxValues();
    const bool ok = vol->write( filename, ascii != 0, external != 0 );
    delete vol;
    return ok ? 0 : 1;
}

struct ref_read_volume
{
    void* values;                 // malloc'ed copy of the value array
    uint64_t nbytes;
    int32_t dtype;                // kvs::Type::TypeID order
    uint32_t dims[3];
    uint32_t veclen;
    int32_t grid_type;            // kvs::StructuredVolumeObject::GridType
    double min_value, max_value;
    float min_obj[3], max_obj[3];
};

// StructuredVolumeObject::read( filename ) (KVSML): what the reference's
// reader makes of a file.
int ref_read_kvsml_volume( const char* filename, ref_read_volume* out )
{
    std::memset( out, 0, sizeof( *out ) );
    kvs::StructuredVolumeObject vol;
    if ( !vol.read( filename ) ) return 1;
    const auto& a = vol.values();
    out->nbytes = a.byteSize();
    out->values = std::malloc( out->nbytes ? out->nbytes : 1 );
    std::memcpy( out->values, a.data(), out->nbytes );
    out->dtype = int32_t( a.typeID() ) - int32_t( kvs::Type::TypeInt8 );
    for ( int k = 0; k < 3; ++k ) out->dims[k] = vol.resolution()[k];
    out->veclen = uint32_t( vol.veclen() );
    out->grid_type = int32_t( vol.gridType() );
    out->min_value = vol.minValue();
    out->max_value = vol.maxValue();
    for ( int k = 0; k < 3; ++k )
    {
        out->min_obj[k] = vol.minObjectCoord()[k];
        out->max_obj[k] = vol.maxObjectCoord()[k];
    }
    return 0;
}

void ref_read_volume_free( ref_read_volume* r ) { std::free( r->values ); std::memset( r, 0, sizeof( *r ) ); }

// AVSField::write of a uniform 3-D field (byte / short / integer / float /
// double values, kvs::Type::TypeID order as dtype), with min_ext / max_ext
// when has_ext.
int ref_write_avs_volume( const void* values, int32_t dtype, const uint32_t dims[3], int32_t veclen, int32_t has_ext,
                          const float ext[6], const char* filename )
{
    kvs::AVSField f;
    const size_t n = size_t( dims[0] ) * dims[1] * dims[2] * size_t( veclen );
    auto set = [&] ( auto tag, kvs::AVSField::DataType type, int bits, bool sign )
    {
        using T = decltype( tag );
        kvs::ValueArray<T> a( n );
        std::memcpy( a.data(), values, n * sizeof( T ) );
        f.setValues( kvs::AnyValueArray( a ) );
        f.setDataType( type );
        f.setBits( bits );
        f.setSigned( sign );
    };
    switch ( dtype )
    {
    case 4: set( kvs::UInt8(), kvs::AVSField::Byte, 8, false ); break;
    case 1: set( kvs::Int16(), kvs::AVSField::Short, 16, true ); break;
    case 2: set( kvs::Int32(), kvs::AVSField::Integer, 32, true ); break;
    case 8: set( kvs::Real32(), kvs::AVSField::Float, 32, true ); break;
    case 9: set( kvs::Real64(), kvs::AVSField::Double, 64, true ); break;
    default: return 2;
    }
    f.setVeclen( veclen );
    f.setNSpace( 3 );
    f.setNDim( 3 );
    f.setDim( kvs::Vec3ui( dims[0], dims[1], dims[2] ) );
    f.setFieldType( kvs::AVSField::Uniform );
    if ( has_ext )
    {
        f.setMinExt( kvs::Vec3( ext[0], ext[1], ext[2] ) );
        f.setMaxExt( kvs::Vec3( ext[3], ext[4], ext[5] ) );
    }
    return f.write( filename ) ? 0 : 1;
}

// StructuredVolumeImporter( filename ) for an AVS field: what the reference's
// importer makes of the file (ext: its min / max external coordinates).  Its
// success flag is not set on the AVS path (ImporterBase.h:26 leaves it
// uninitialised and import( const AVSField* ) only ever clears it), so a
// refused or unreadable file is told by the values it did not set.
int ref_read_avs_volume( const char* filename, ref_read_volume* out, float ext[6] )
{
    std::memset( out, 0, sizeof( *out ) );
    kvs::StructuredVolumeImporter vol( filename );
    if ( vol.values().size() == 0 ) return 1;
    const auto& a = vol.values();
    out->nbytes = a.byteSize();
    out->values = std::malloc( out->nbytes ? out->nbytes : 1 );
    std::memcpy( out->values, a.data(), out->nbytes );
    out->dtype = int32_t( a.typeID() ) - int32_t( kvs::Type::TypeInt8 );
    for ( int k = 0; k < 3; ++k ) out->dims[k] = vol.resolution()[k];
    out->veclen = uint32_t( vol.veclen() );
    out->grid_type = int32_t( vol.gridType() );
    out->min_value = vol.minValue();
    out->max_value = vol.maxValue();
    for ( int k = 0; k < 3; ++k )
    {
        out->min_obj[k] = vol.minObjectCoord()[k];
        out->max_obj[k] = vol.maxObjectCoord()[k];
        ext[k] = vol.minExternalCoord()[k];
        ext[3 + k] = vol.maxExternalCoord()[k];
    }
    return 0;
}

// PolygonExporter<kvs::Stl>( polygon ) and Stl::write( filename ) for a
// triangle polygon with the given arrays and normal type.
int ref_export_stl( const float* coords, uint64_t ncoords, const uint32_t* conn, uint64_t nconn,
                    const float* normals, uint64_t nnormals, int normal_type, int ascii, const char* filename )
{
    kvs::PolygonObject p;
    kvs::ValueArray<kvs::Real32> c( ncoords ); std::memcpy( c.data(), coords, ncoords * 4 );
    kvs::ValueArray<kvs::UInt32> k( nconn ); if ( nconn ) std::memcpy( k.data(), conn, nconn * 4 );
    kvs::ValueArray<kvs::Real32> n( nnormals ); if ( nnormals ) std::memcpy( n.data(), normals, nnormals * 4 );
    p.setCoords( c );
    if ( nconn ) p.setConnections( k );
    p.setNormals( n );
    p.setPolygonType( kvs::PolygonObject::Triangle );
    p.setNormalType( kvs::PolygonObject::NormalType( normal_type ) );
    p.setColorType( kvs::PolygonObject::PolygonColor );
    p.setColor( kvs::RGBColor( 255, 255, 255 ) );
    kvs::PolygonExporter<kvs::Stl> e( &p );
    e.setFileType( ascii ? kvs::Stl::Ascii : kvs::Stl::Binary );
    return e.write( filename ) ? 0 : 1;
}

// PolygonExporter<kvs::KVSMLPolygonObject>( polygon ), setWritingDataType( type )
// and write( filename ) for a triangle polygon with the given arrays and types
// (ncolors bytes of colour, opacity < 0: no opacity).
int ref_export_kvsml_polygon( const float* coords, uint64_t ncoords, const uint32_t* conn, uint64_t nconn,
                              const float* normals, uint64_t nnormals, const uint8_t* colors, uint64_t ncolors,
                              int opacity, int color_type, int normal_type, int writing_type, const char* filename )
{
    kvs::PolygonObject p;
    kvs::ValueArray<kvs::Real32> c( ncoords ); if ( ncoords ) std::memcpy( c.data(), coords, ncoords * 4 );
    p.setCoords( c );
    if ( nconn ) { kvs::ValueArray<kvs::UInt32> k( nconn ); std::memcpy( k.data(), conn, nconn * 4 ); p.setConnections( k ); }
    if ( nnormals ) { kvs::ValueArray<kvs::Real32> n( nnormals ); std::memcpy( n.data(), normals, nnormals * 4 ); p.setNormals( n ); }
    if ( ncolors ) { kvs::ValueArray<kvs::UInt8> k( ncolors ); std::memcpy( k.data(), colors, ncolors ); p.setColors( k ); }
    if ( opacity >= 0 ) p.setOpacity( kvs::UInt8( opacity ) );
    p.setPolygonType( kvs::PolygonObject::Triangle );
    p.setColorType( kvs::PolygonObject::ColorType( color_type ) );
    p.setNormalType( kvs::PolygonObject::NormalType( normal_type ) );
    kvs::PolygonExporter<kvs::KVSMLPolygonObject> e( &p );
    e.setWritingDataType( kvs::KVSMLPolygonObject::WritingDataType( writing_type ) );
    return e.write( filename ) ? 0 : 1;
}

// PolygonExporter<kvs::Ply>( polygon ), setFileType and write( filename ) for
// a triangle polygon without connections: with connections the reference's
// writer stores face indices through a null pointer (Ply.cpp:475-478), so
// this refuses them (returns 2).
int ref_export_ply( const float* coords, uint64_t ncoords, uint64_t nconn, const float* normals, uint64_t nnormals,
                    const uint8_t* colors, uint64_t ncolors, int color_type, int normal_type, int binary,
                    const char* filename )
{
    if ( nconn ) return 2;
    kvs::PolygonObject p;
    kvs::ValueArray<kvs::Real32> c( ncoords ); if ( ncoords ) std::memcpy( c.data(), coords, ncoords * 4 );
    p.setCoords( c );
    if ( nnormals ) { kvs::ValueArray<kvs::Real32> n( nnormals ); std::memcpy( n.data(), normals, nnormals * 4 ); p.setNormals( n ); }
    if ( ncolors ) { kvs::ValueArray<kvs::UInt8> k( ncolors ); std::memcpy( k.data(), colors, ncolors ); p.setColors( k ); }
    p.setPolygonType( kvs::PolygonObject::Triangle );
    p.setColorType( kvs::PolygonObject::ColorType( color_type ) );
    p.setNormalType( kvs::PolygonObject::NormalType( normal_type ) );
    kvs::PolygonExporter<kvs::Ply> e( &p );
    e.setFileType( binary ? kvs::Ply::Binary : kvs::Ply::Ascii );
    return e.write( filename ) ? 0 : 1;
}

// Ply::read( filename ): the coordinates, colours, normals and connections the
// reference's reader makes of a PLY file (free with ref_mesh_free).
int ref_read_ply( const char* filename, ref_mesh* out )
{
    std::memset( out, 0, sizeof( *out ) );
    kvs::Ply ply;
    if ( !ply.read( filename ) ) return 1;
    out->coords = static_cast<float*>( CopyOut( ply.coords(), &out->n_coords ) );
    out->connections = static_cast<uint32_t*>( CopyOut( ply.connections(), &out->n_connections ) );
    out->normals = static_cast<float*>( CopyOut( ply.normals(), &out->n_normals ) );
    out->colors = static_cast<uint8_t*>( CopyOut( ply.colors(), &out->n_colors ) );
    out->success = ply.hasColors() ? 1 : 0;
    return 0;
}

// MarchingCubesTable::TriangleID and VertexID, for pinning transcriptions.
void ref_mc_tables( int32_t tri[256 * 16], int32_t vert[12 * 2 * 3] )
{
    std::memcpy( tri, kvs::MarchingCubesTable::TriangleID, sizeof( int32_t ) * 256 * 16 );
    std::memcpy( vert, kvs::MarchingCubesTable::VertexID, sizeof( int32_t ) * 12 * 2 * 3 );
}

} // extern "C"
