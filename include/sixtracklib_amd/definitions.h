/* sixtracklib_amd/definitions.h
 *
 * Scalar types, the C name prefix, object type ids, status codes and the
 * compile-time tracking configuration shared by the host library, the HIP
 * kernels and the CPU oracle.
 *
 * Names and values follow the reference so that code written against the
 * reference C API compiles against this library unchanged:
 *   - NS(x) expands to st_x (sixtracklib/common/generated/namespace.h.template,
 *     default prefix "st_").
 *   - type ids: sixtracklib/common/internal/objects_type_id.h:20-41
 *   - status codes: sixtracklib/common/track/definitions.h:12-18,24
 *   - IO flags: sixtracklib/common/track/definitions.h:43-71
 *   - aperture configuration defaults: Settings.cmake.default and
 *     sixtracklib/common/generated/config.h.template (global aperture
 *     "always" at every Drift/DriftExact, X/Y limit 1.0 m).
 */
#ifndef SIXTRACKLIB_AMD_DEFINITIONS_H__
#define SIXTRACKLIB_AMD_DEFINITIONS_H__

#include <stdint.h>
#include <stddef.h>
#if !defined( __cplusplus )
    #include <stdbool.h>
#endif

#if !defined( NS )
    #define NS( name ) st_##name
#endif

#if defined( __cplusplus )
    #define ST_EXTERN_C extern "C"
#else
    #define ST_EXTERN_C extern
#endif

#if defined( __GNUC__ )
    #define ST_API __attribute__((visibility("default")))
#else
    #define ST_API
#endif

typedef uint64_t NS(buffer_size_t);
typedef uint64_t NS(buffer_addr_t);
typedef int64_t  NS(buffer_addr_diff_t);
typedef uint64_t NS(object_type_id_t);
typedef double   NS(particle_real_t);
typedef int64_t  NS(particle_index_t);
typedef int64_t  NS(particle_num_elements_t);
typedef int32_t  NS(track_status_t);
typedef int32_t  NS(arch_status_t);
typedef uint64_t NS(arch_size_t);
typedef uint16_t NS(track_job_io_flag_t);
typedef NS(track_job_io_flag_t) NS(track_job_collect_flag_t);
typedef NS(track_job_io_flag_t) NS(track_job_push_flag_t);
typedef uint64_t NS(track_job_size_t);
typedef int64_t  NS(track_job_type_t);
typedef int64_t  NS(multipole_order_t);
typedef int64_t  NS(be_monitor_turn_t);
typedef int64_t  NS(be_monitor_index_t);
typedef int64_t  NS(be_monitor_flag_t);
typedef uint64_t NS(be_monitor_addr_t);

#define ST_TRACK_SUCCESS               ( 0 )
#define ST_TRACK_STATUS_GENERAL_FAILURE ( -1 )
#define ST_ARCH_STATUS_SUCCESS         ( 0 )
#define ST_ARCH_STATUS_GENERAL_FAILURE ( -1 )

#define ST_TRACK_JOB_IO_NONE            ( 0x00 )
#define ST_TRACK_JOB_IO_PARTICLES       ( 0x01 )
#define ST_TRACK_JOB_IO_BEAM_ELEMENTS   ( 0x02 )
#define ST_TRACK_JOB_IO_OUTPUT          ( 0x04 )
#define ST_TRACK_JOB_IO_DEBUG_REGISTER  ( 0x08 )
#define ST_TRACK_JOB_IO_PARTICLES_ADDR  ( 0x10 )
#define ST_TRACK_JOB_COLLECT_ALL        ( 0x1f )
#define ST_TRACK_JOB_PUSH_ALL           ( 0x0f )
#define ST_TRACK_JOB_COLLECT_DEFAULT_FLAGS ( 0x05 )
#define ST_TRACK_JOB_PUSH_DEFAULT_FLAGS    ( 0x02 )

/* Object type ids (values identical to the reference). */
#define ST_OBJECT_TYPE_NONE                  ( 0 )
#define ST_OBJECT_TYPE_PARTICLE              ( 1 )
#define ST_OBJECT_TYPE_DRIFT                 ( 2 )
#define ST_OBJECT_TYPE_DRIFT_EXACT           ( 3 )
#define ST_OBJECT_TYPE_MULTIPOLE             ( 4 )
#define ST_OBJECT_TYPE_CAVITY                ( 5 )
#define ST_OBJECT_TYPE_XYSHIFT               ( 6 )
#define ST_OBJECT_TYPE_SROTATION             ( 7 )
#define ST_OBJECT_TYPE_BEAM_BEAM_4D          ( 8 )
#define ST_OBJECT_TYPE_BEAM_BEAM_6D          ( 9 )
#define ST_OBJECT_TYPE_BEAM_MONITOR          ( 10 )
#define ST_OBJECT_TYPE_LIMIT_RECT            ( 11 )
#define ST_OBJECT_TYPE_LIMIT_ELLIPSE         ( 12 )
#define ST_OBJECT_TYPE_SC_COASTING           ( 13 )
#define ST_OBJECT_TYPE_SC_QGAUSSIAN_PROF     ( 14 )
#define ST_OBJECT_TYPE_LIMIT_RECT_ELLIPSE    ( 16 )
#define ST_OBJECT_TYPE_DIPEDGE               ( 24 )
#define ST_OBJECT_TYPE_RF_MULTIPOLE          ( 256 )
#define ST_OBJECT_TYPE_SC_INTERPOLATED_PROF  ( 257 )
#define ST_OBJECT_TYPE_LINE_DENSITY_PROF_DATA ( 258 )
#define ST_OBJECT_TYPE_LIN_DECOUPLED_ROTATION ( 259 )
#define ST_OBJECT_TYPE_TRICUB_DATA           ( 260 )
#define ST_OBJECT_TYPE_TRICUB                ( 261 )
#define ST_OBJECT_TYPE_PARTICLES_ADDR        ( 512 )
#define ST_OBJECT_TYPE_LINE                  ( 1024 )
#define ST_OBJECT_TYPE_ELEM_BY_ELEM_CONF     ( 65534 )
#define ST_OBJECT_TYPE_NODE_ID               ( 65535 )
#define ST_OBJECT_TYPE_INVALID               ( 0x7fffffff )

/* Compile-time tracking configuration = the reference defaults. */
#define ST_APERTURE_X_LIMIT ( 1.0 )
#define ST_APERTURE_Y_LIMIT ( 1.0 )

/* Physical constants, literal-identical to the reference
 * (sixtracklib/common/constants.h:10-36,
 *  sixtracklib/common/internal/physics_constants.h:28-69). */
#define ST_C_LIGHT   ( 299792458.0 )
#define ST_EPSILON_0 ( 8.854187817620e-12 )
#define ST_PI        ( 3.1415926535897932384626433832795028841971693993751 )
#define ST_DEG2RAD   ( 0.0174532925199432957692369076848861271344287188854 )
#define ST_SQRT_PI   ( 1.7724538509055160272981674833411451827975494561224 )
#define ST_QELEM     ( 1.60217662e-19 )
#define ST_CHARGE0_SI ( 1.602176634e-19 )

#endif /* SIXTRACKLIB_AMD_DEFINITIONS_H__ */
