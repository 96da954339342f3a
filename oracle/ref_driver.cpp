// C entry points around the reference library's own dip::Gauss, for the test
// suite and the CPU baseline only.  Compiled against the reference sources by
// oracle/ref.mk into oracle/_ref/libdipref.so.  The product never links this.
//
// Data-type codes are the ones of include/dipamd.h (DIPAMD_DT_*).
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "diplib.h"
#include "diplib/linear.h"
#include "diplib/generation.h"
#include "diplib/multithreading.h"
#include "diplib/analysis.h"

namespace {

thread_local std::string g_err;

dip::DataType ToDip( int code ) {
   switch( code ) {
      case 0: return dip::DT_UINT8;
      case 1: return dip::DT_UINT16;
      case 2: return dip::DT_UINT32;
      case 3: return dip::DT_SINT8;
      case 4: return dip::DT_SINT16;
      case 5: return dip::DT_SINT32;
      case 6: return dip::DT_SFLOAT;
      case 7: return dip::DT_DFLOAT;
      case 8: return dip::DT_BIN;
      case 9: return dip::DT_SCOMPLEX;
      case 10: return dip::DT_DCOMPLEX;
      default: throw std::runtime_error( "unknown dtype code" );
   }
}

dip::Image Wrap( void const* ptr, int dtype, int ndim, int64_t const* sizes ) {
   dip::UnsignedArray sz( static_cast< dip::uint >( ndim ));
   for( int ii = 0; ii < ndim; ++ii ) {
      sz[ static_cast< dip::uint >( ii ) ] = static_cast< dip::uint >( sizes[ ii ] );
   }
   void* p = const_cast< void* >( ptr );
   return dip::Image( dip::NonOwnedRefToDataSegment( p ), p, ToDip( dtype ), sz );
}

} // namespace

extern "C" {

const char* dipref_last_error() { return g_err.c_str(); }

// dip::Gauss( in, out, sigmas, orders, method, bc, truncation ) on host buffers with
// normal strides (dim 0 contiguous).  `out` is wrapped as a protected image of
// `out_dtype`, so the reference writes into it in place.  `bcs` holds `nbc` C strings
// (0, 1 or ndim of them).  nthreads == 0 leaves the reference's default.
int dipref_gauss( void const* in, int in_dtype, void* out, int out_dtype, int ndim, int64_t const* sizes,
                  double const* sigmas, int nsigmas, int const* orders, int norders,
                  char const* const* bcs, int nbc, double truncation, char const* method, int nthreads ) {
   try {
      dip::Image imgIn = Wrap( in, in_dtype, ndim, sizes );
      dip::Image imgOut = Wrap( out, out_dtype, ndim, sizes );
      imgOut.Protect();
      dip::FloatArray s( static_cast< dip::uint >( nsigmas ));
      for( int ii = 0; ii < nsigmas; ++ii ) { s[ static_cast< dip::uint >( ii ) ] = sigmas[ ii ]; }
      dip::UnsignedArray o( static_cast< dip::uint >( norders ));
      for( int ii = 0; ii < norders; ++ii ) { o[ static_cast< dip::uint >( ii ) ] = static_cast< dip::uint >( orders[ ii ] ); }
      dip::StringArray bc;
      for( int ii = 0; ii < nbc; ++ii ) { bc.emplace_back( bcs[ ii ] ); }
      if( nthreads > 0 ) {
         dip::SetNumberOfThreads( static_cast< dip::uint >( nthreads ));
      }
      dip::Gauss( imgIn, imgOut, s, o, method, bc, truncation );
      return 0;
   } catch( std::exception const& e ) {
      g_err = e.what();
      return 1;
   }
}

// dip::GaussIIR( in, out, sigmas, orders, bc, filterOrder, designMethod, truncation ):
// as dipref_gauss; `filter_orders` holds `nfo` values (0 = the default per order).
int dipref_gauss_iir( void const* in, int in_dtype, void* out, int out_dtype, int ndim, int64_t const* sizes,
                      double const* sigmas, int nsigmas, int const* orders, int norders,
                      char const* const* bcs, int nbc, int const* filter_orders, int nfo,
                      char const* method, double truncation, int nthreads ) {
   try {
      dip::Image imgIn = Wrap( in, in_dtype, ndim, sizes );
      dip::Image imgOut = Wrap( out, out_dtype, ndim, sizes );
      imgOut.Protect();
      dip::FloatArray s( static_cast< dip::uint >( nsigmas ));
      for( int ii = 0; ii < nsigmas; ++ii ) { s[ static_cast< dip::uint >( ii ) ] = sigmas[ ii ]; }
      dip::UnsignedArray o( static_cast< dip::uint >( norders ));
      for( int ii = 0; ii < norders; ++ii ) { o[ static_cast< dip::uint >( ii ) ] = static_cast< dip::uint >( orders[ ii ] ); }
      dip::UnsignedArray fo( static_cast< dip::uint >( nfo ));
      for( int ii = 0; ii < nfo; ++ii ) { fo[ static_cast< dip::uint >( ii ) ] = static_cast< dip::uint >( filter_orders[ ii ] ); }
      dip::StringArray bc;
      for( int ii = 0; ii < nbc; ++ii ) { bc.emplace_back( bcs[ ii ] ); }
      if( nthreads > 0 ) {
         dip::SetNumberOfThreads( static_cast< dip::uint >( nthreads ));
      }
      dip::GaussIIR( imgIn, imgOut, s, o, bc, fo, method, truncation );
      return 0;
   } catch( std::exception const& e ) {
      g_err = e.what();
      return 1;
   }
}

// dip::GaussFT( in, out, sigmas, orders, truncation, "spatial", "spatial", bc ): as
// dipref_gauss; nbc == 0 passes an empty boundary condition array (no extension).
int dipref_gauss_ft( void const* in, int in_dtype, void* out, int out_dtype, int ndim, int64_t const* sizes,
                     double const* sigmas, int nsigmas, int const* orders, int norders,
                     char const* const* bcs, int nbc, double truncation ) {
   try {
      dip::Image imgIn = Wrap( in, in_dtype, ndim, sizes );
      dip::Image imgOut = Wrap( out, out_dtype, ndim, sizes );
      imgOut.Protect();
      dip::FloatArray s( static_cast< dip::uint >( nsigmas ));
      for( int ii = 0; ii < nsigmas; ++ii ) { s[ static_cast< dip::uint >( ii ) ] = sigmas[ ii ]; }
      dip::UnsignedArray o( static_cast< dip::uint >( norders ));
      for( int ii = 0; ii < norders; ++ii ) { o[ static_cast< dip::uint >( ii ) ] = static_cast< dip::uint >( orders[ ii ] ); }
      dip::StringArray bc;
      for( int ii = 0; ii < nbc; ++ii ) { bc.emplace_back( bcs[ ii ] ); }
      dip::GaussFT( imgIn, imgOut, s, o, truncation, "spatial", "spatial", bc );
      return 0;
   } catch( std::exception const& e ) {
      g_err = e.what();
      return 1;
   }
}

// dip::GaussFT( in, out, sigmas, orders, truncation, inRepr, outRepr, {} ) with the
// domains named ("spatial" / "frequency"); the output is written into `out` (protected,
// so of out_dtype and the input's sizes: no boundary extension).
int dipref_gauss_ft_repr( void const* in, int in_dtype, void* out, int out_dtype, int ndim, int64_t const* sizes,
                          double const* sigmas, int nsigmas, int const* orders, int norders, double truncation,
                          char const* in_repr, char const* out_repr ) {
   try {
      dip::Image imgIn = Wrap( in, in_dtype, ndim, sizes );
      dip::Image imgOut = Wrap( out, out_dtype, ndim, sizes );
      imgOut.Protect();
      dip::FloatArray s( static_cast< dip::uint >( nsigmas ));
      for( int ii = 0; ii < nsigmas; ++ii ) { s[ static_cast< dip::uint >( ii ) ] = sigmas[ ii ]; }
      dip::UnsignedArray o( static_cast< dip::uint >( norders ));
      for( int ii = 0; ii < norders; ++ii ) { o[ static_cast< dip::uint >( ii ) ] = static_cast< dip::uint >( orders[ ii ] ); }
      dip::GaussFT( imgIn, imgOut, s, o, truncation, in_repr, out_repr, {} );
      return 0;
   } catch( std::exception const& e ) {
      g_err = e.what();
      return 1;
   }
}

// In-place form: `buf` is both the input and the output (same dtype, which must be
// sfloat or dfloat).
int dipref_gauss_inplace( void* buf, int dtype, int ndim, int64_t const* sizes, double const* sigmas, int nsigmas,
                          int const* orders, int norders, char const* bc, double truncation ) {
   try {
      dip::Image img = Wrap( buf, dtype, ndim, sizes );
      img.Protect();
      dip::FloatArray s( static_cast< dip::uint >( nsigmas ));
      for( int ii = 0; ii < nsigmas; ++ii ) { s[ static_cast< dip::uint >( ii ) ] = sigmas[ ii ]; }
      dip::UnsignedArray o( static_cast< dip::uint >( norders ));
      for( int ii = 0; ii < norders; ++ii ) { o[ static_cast< dip::uint >( ii ) ] = static_cast< dip::uint >( orders[ ii ] ); }
      dip::Gauss( img, img, s, o, "FIR", { bc }, truncation );
      return 0;
   } catch( std::exception const& e ) {
      g_err = e.what();
      return 1;
   }
}

// dip::MakeHalfGaussian: writes up to `cap` taps (x = 0 last) and returns the count,
// or -1 on error.  is_double selects the dfloat truncation limit.
int dipref_half_gaussian( double sigma, int order, double truncation, int is_double, double* taps, int cap ) {
   try {
      std::vector< dip::dfloat > f = dip::MakeHalfGaussian( sigma, static_cast< dip::uint >( order ), truncation,
                                                            is_double ? dip::DT_DFLOAT : dip::DT_SFLOAT );
      int n = static_cast< int >( f.size() );
      if( n > cap ) { g_err = "capacity"; return -1; }
      std::memcpy( taps, f.data(), f.size() * sizeof( double ));
      return n;
   } catch( std::exception const& e ) {
      g_err = e.what();
      return -1;
   }
}

// dip::SeparableConvolution with one 1-D filter per dimension (or one for all), given as
// full tap arrays with a symmetry string each ("general", "even", "odd", "conj", "d-even",
// "d-odd", "d-conj"), an origin each (`origins` NULL: -1, the centre) and a complex flag
// each (`is_complex` NULL: all real; a complex filter's array holds re, im pairs and
// `ntaps` counts doubles, as OneDimensionalFilter::filter does).
int dipref_separable( void const* in, int in_dtype, void* out, int out_dtype, int ndim, int64_t const* sizes,
                      double const* const* taps, int const* ntaps, char const* const* symmetry, int nfilters,
                      char const* const* bcs, int nbc, unsigned char const* process, int const* origins,
                      int const* is_complex ) {
   try {
      dip::Image imgIn = Wrap( in, in_dtype, ndim, sizes );
      dip::Image imgOut = Wrap( out, out_dtype, ndim, sizes );
      imgOut.Protect();
      dip::OneDimensionalFilterArray fa( static_cast< dip::uint >( nfilters ));
      for( int ii = 0; ii < nfilters; ++ii ) {
         auto& f = fa[ static_cast< dip::uint >( ii ) ];
         f.filter.assign( taps[ ii ], taps[ ii ] + ntaps[ ii ] );
         f.symmetry = symmetry[ ii ];
         if( origins ) { f.origin = origins[ ii ]; }
         if( is_complex ) { f.isComplex = is_complex[ ii ] != 0; }
      }
      dip::StringArray bc;
      for( int ii = 0; ii < nbc; ++ii ) { bc.emplace_back( bcs[ ii ] ); }
      dip::BooleanArray proc;
      if( process ) {
         proc.resize( static_cast< dip::uint >( ndim ));
         for( int ii = 0; ii < ndim; ++ii ) { proc[ static_cast< dip::uint >( ii ) ] = process[ ii ] != 0; }
      }
      dip::SeparableConvolution( imgIn, imgOut, fa, bc, proc );
      return 0;
   } catch( std::exception const& e ) {
      g_err = e.what();
      return 1;
   }
}

// The derivative family on dip::Gauss (derivative.cpp:142-521): `which` is one of
// "gradient", "gradmag", "hessian", "laplace", "dgg", "laplace+dgg", "laplace-dgg", or
// "structure:<tensor sigma>" (dip::StructureTensor, no mask, structure_tensor.cpp:32-57).
// The result's tensor elements are copied, in the reference's element order, into
// `out` as `cap` consecutive images of the input's sizes and of type `out_dtype`
// (sfloat or dfloat); the element count is returned in *nelem.
int dipref_derivative( char const* which, void const* in, int in_dtype, void* out, int out_dtype, int cap,
                       int* nelem, int ndim, int64_t const* sizes, double const* sigmas, int nsigmas,
                       char const* const* bcs, int nbc, char const* method, double truncation ) {
   try {
      dip::Image imgIn = Wrap( in, in_dtype, ndim, sizes );
      dip::FloatArray s( static_cast< dip::uint >( nsigmas ));
      for( int ii = 0; ii < nsigmas; ++ii ) { s[ static_cast< dip::uint >( ii ) ] = sigmas[ ii ]; }
      dip::StringArray bc;
      for( int ii = 0; ii < nbc; ++ii ) { bc.emplace_back( bcs[ ii ] ); }
      std::string w( which );
      dip::Image res;
      if( w == "gradient" ) { dip::Gradient( imgIn, res, s, method, bc, {}, truncation ); }
      else if( w == "gradmag" ) { dip::GradientMagnitude( imgIn, res, s, method, bc, {}, truncation ); }
      else if( w == "hessian" ) { dip::Hessian( imgIn, res, s, method, bc, {}, truncation ); }
      else if( w == "laplace" ) { dip::Laplace( imgIn, res, s, method, bc, {}, truncation ); }
      else if( w == "dgg" ) { dip::Dgg( imgIn, res, s, method, bc, {}, truncation ); }
      else if( w == "laplace+dgg" ) { dip::LaplacePlusDgg( imgIn, res, s, method, bc, {}, truncation ); }
      else if( w == "laplace-dgg" ) { dip::LaplaceMinusDgg( imgIn, res, s, method, bc, {}, truncation ); }
      else if( w.substr( 0, 10 ) == "structure:" ) {  // "structure:<tensor sigma>"
         dip::StructureTensor( imgIn, {}, res, s, { std::stod( w.substr( 10 )) }, method, bc, truncation );
      }
      else { throw std::runtime_error( "unknown function" ); }
      int n = static_cast< int >( res.TensorElements() );
      *nelem = n;
      if( n > cap ) { g_err = "capacity"; return 1; }
      dip::uint npx = 1;
      for( int ii = 0; ii < ndim; ++ii ) { npx *= static_cast< dip::uint >( sizes[ ii ] ); }
      std::size_t es = out_dtype == 7 ? 8 : 4;
      for( int ii = 0; ii < n; ++ii ) {
         dip::Image dst = Wrap( static_cast< char* >( out ) + static_cast< std::size_t >( ii ) * npx * es, out_dtype,
                                ndim, sizes );
         dst.Protect();
         dst.Copy( res[ static_cast< dip::uint >( ii ) ] );
      }
      return 0;
   } catch( std::exception const& e ) {
      g_err = e.what();
      return 1;
   }
}

// The masked forms: "nc" dip::NormalizedConvolution, "ndc:<dim>"
// dip::NormalizedDifferentialConvolution along <dim>, "structure:<tensor sigma>"
// dip::StructureTensor with the mask; as dipref_derivative otherwise.
int dipref_derivative_masked( char const* which, void const* in, int in_dtype, void const* mask, int mask_dtype,
                              void* out, int out_dtype, int cap, int* nelem, int ndim, int64_t const* sizes,
                              double const* sigmas, int nsigmas, char const* const* bcs, int nbc, char const* method,
                              double truncation ) {
   try {
      dip::Image imgIn = Wrap( in, in_dtype, ndim, sizes );
      dip::Image imgMask = Wrap( mask, mask_dtype, ndim, sizes );
      dip::FloatArray s( static_cast< dip::uint >( nsigmas ));
      for( int ii = 0; ii < nsigmas; ++ii ) { s[ static_cast< dip::uint >( ii ) ] = sigmas[ ii ]; }
      dip::StringArray bc;
      for( int ii = 0; ii < nbc; ++ii ) { bc.emplace_back( bcs[ ii ] ); }
      std::string w( which );
      dip::Image res;
      if( w == "nc" ) { dip::NormalizedConvolution( imgIn, imgMask, res, s, method, bc, truncation ); }
      else if( w.substr( 0, 4 ) == "ndc:" ) {
         dip::NormalizedDifferentialConvolution( imgIn, imgMask, res, std::stoul( w.substr( 4 )), s, method, bc, truncation );
      }
      else if( w.substr( 0, 10 ) == "structure:" ) {
         dip::StructureTensor( imgIn, imgMask, res, s, { std::stod( w.substr( 10 )) }, method, bc, truncation );
      }
      else { throw std::runtime_error( "unknown function" ); }
      int n = static_cast< int >( res.TensorElements() );
      *nelem = n;
      if( n > cap ) { g_err = "capacity"; return 1; }
      dip::uint npx = 1;
      for( int ii = 0; ii < ndim; ++ii ) { npx *= static_cast< dip::uint >( sizes[ ii ] ); }
      std::size_t es = out_dtype == 7 ? 8 : 4;
      for( int ii = 0; ii < n; ++ii ) {
         dip::Image dst = Wrap( static_cast< char* >( out ) + static_cast< std::size_t >( ii ) * npx * es, out_dtype,
                                ndim, sizes );
         dst.Protect();
         dst.Copy( res[ static_cast< dip::uint >( ii ) ] );
      }
      return 0;
   } catch( std::exception const& e ) {
      g_err = e.what();
      return 1;
   }
}

int dipref_max_threads() { return static_cast< int >( dip::GetNumberOfThreads() ); }

} // extern "C"
