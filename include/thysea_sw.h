// thysea_sw.h -- the reference's boundary function, with its exact signature, on the
// MI355X library: a caller written for the reference's GPU part links against
// libthysea_sw.so instead and keeps its own setup code.
//
//   reference: extern "C" int shallowWater(... 53 arguments ...)
//              (src/GPU/ShallowWater.cu:182-190, declared src/ShallowWater.cxx:19-27,
//               called src/ShallowWater.cxx:146-151)
//
// CUDA's double2 and the reference's tipoDatosSubmalla (src/Constantes.hxx:32-38) are
// replaced by plain structs of the same layout, so the reference's caller, compiled with
// its own headers, passes the same bytes.  As in the reference, prefijo is a
// std::string passed by value (the one argument that is not plain C).
#pragma once
#include <stdint.h>

#include <string>

// layout of CUDA's double2 (16 bytes, 16-byte aligned)
struct alignas(16) thy_double2 {
    double x, y;
};

// layout of tipoDatosSubmalla: per-row (area, cos phi), cell widths [ny+1], heights [ny],
// longitudes [nx], latitudes [ny] (built by src/Problema.cxx:431-448)
struct thy_submalla {
    thy_double2 *areaYCosPhi;
    double *anchoVolumenes;
    double *altoVolumenes;
    double *longitud;
    double *latitud;
};

// Same arguments, meaning and return codes as the reference (0 ok, 1 not enough GPU
// memory, 2 not enough CPU memory; further codes as thy_status in include/thysea.h).
// Inputs are non-dimensional as cargarDatosProblema leaves them
// (src/Problema.cxx:327-473); defTime in seconds; *tiempo receives the loop's wall time.
// Writes <prefijo>.nc when tiempoGuardarNetCDF >= 0 and <prefijo>_ts.nc when
// leer_fichero_puntos == 1.  Only one level is supported (numNiveles == 1), as in the
// reference.  The host arrays are read only (the reference reuses
// datosVolumenesNivel_1/_2 as scratch for gauge data; this function does not).
extern "C" int shallowWater(int numNiveles, int okada_flag, int numFaults, int crop_flag, double crop_value,
                            double *LON_C, double *LAT_C, double *DEPTH_C, double *FAULT_L, double *FAULT_W,
                            double *STRIKE, double *DIP, double *RAKE, double *SLIP, double *defTime,
                            thy_double2 *datosVolumenesNivel_1, thy_double2 *datosVolumenesNivel_2,
                            thy_submalla datosNivel, int leer_fichero_puntos, int numPuntosGuardar,
                            int *posicionesVolumenesGuardado, double *lonPuntos, double *latPuntos,
                            int numVolxNivel0, int numVolyNivel0, int64_t numVolumenesNivel, double Hmin,
                            char *nombre_bati, std::string prefijo, double borde_sup, double borde_inf,
                            double borde_izq, double borde_der, int tam_spongeSup, int tam_spongeInf,
                            int tam_spongeIzq, int tam_spongeDer, double tiempo_tot, double tiempoGuardarNetCDF,
                            double tiempoGuardarSeries, double CFL, double mf0, double vmax, double epsilon_h,
                            double hpos, double cvis, double dif_at, double L, double H, double Q, double T,
                            char *version, double *tiempo);
