// shallow_water.cpp -- the reference's boundary function shallowWater
// (src/GPU/ShallowWater.cu:182-500) with its exact signature: the arguments become one
// run description and run_problem does the work on the C-ABI of libthysea.so.
#include "../../include/thysea_sw.h"

#include <cmath>
#include <new>

#include "run.h"

using thyhost::Problem;

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
                            char *version, double *tiempo) {
    if (numNiveles != 1 || numVolxNivel0 < 2 || numVolyNivel0 < 2 ||
        numVolumenesNivel != (int64_t)numVolxNivel0 * numVolyNivel0 || numFaults < 0 ||
        (okada_flag == thyhost::OKADA_STANDARD && numFaults > 0 && !defTime) || !datosVolumenesNivel_1 ||
        !datosVolumenesNivel_2 || !datosNivel.areaYCosPhi || !datosNivel.anchoVolumenes ||
        !datosNivel.altoVolumenes || !datosNivel.longitud || !datosNivel.latitud)
        return THY_ERR_INVALID;
    const int nx = numVolxNivel0, ny = numVolyNivel0;
    Problem p;
    try {
        p.name = nombre_bati ? nombre_bati : "";
        p.prefix = prefijo;
        p.okada_flag = okada_flag;
        if (okada_flag == thyhost::OKADA_STANDARD) {
            for (int k = 0; k < numFaults; k++)
                p.faults.push_back(thy_fault{defTime[k], LON_C[k], LAT_C[k], DEPTH_C[k], FAULT_L[k], FAULT_W[k],
                                             STRIKE[k], DIP[k], RAKE[k], SLIP[k]});
            p.crop_flag = crop_flag;
            p.crop_value = crop_value;
        }
        p.nx = nx;
        p.ny = ny;
        p.lon.assign(datosNivel.longitud, datosNivel.longitud + nx);
        p.lat.assign(datosNivel.latitud, datosNivel.latitud + ny);
        const double *s1 = &datosVolumenesNivel_1[0].x, *s2 = &datosVolumenesNivel_2[0].x;
        p.state1.assign(s1, s1 + 2 * (size_t)numVolumenesNivel);
        p.state2.assign(s2, s2 + 2 * (size_t)numVolumenesNivel);
        p.hmin = Hmin;
        p.area.resize(ny);
        p.cosphi.resize(ny);
        for (int j = 0; j < ny; j++) {
            p.area[j] = datosNivel.areaYCosPhi[j].x;
            p.cosphi[j] = datosNivel.areaYCosPhi[j].y;
        }
        p.ancho.assign(datosNivel.anchoVolumenes, datosNivel.anchoVolumenes + ny + 1);
        p.alto.assign(datosNivel.altoVolumenes, datosNivel.altoVolumenes + ny);
        p.borde_sup = borde_sup;
        p.borde_inf = borde_inf;
        p.borde_izq = borde_izq;
        p.borde_der = borde_der;
        p.sp_sup = tam_spongeSup;
        p.sp_inf = tam_spongeInf;
        p.sp_izq = tam_spongeIzq;
        p.sp_der = tam_spongeDer;
        p.t_end = tiempo_tot;
        p.t_save = tiempoGuardarNetCDF;
        p.t_series = tiempoGuardarSeries;
        p.cfl = CFL;
        p.mf0 = mf0;
        p.vmax = vmax;
        p.eps_h = epsilon_h;
        p.hpos = hpos;
        p.cvis = cvis;
        p.dif_at = dif_at;
        p.L = L;
        p.H = H;
        p.Q = Q;
        p.T = T;
        p.read_points = leer_fichero_puntos;
        if (leer_fichero_puntos == 1 && numPuntosGuardar > 0) {
            p.ppos.assign(posicionesVolumenesGuardado, posicionesVolumenesGuardado + numPuntosGuardar);
            p.plon.assign(lonPuntos, lonPuntos + numPuntosGuardar);
            p.plat.assign(latPuntos, latPuntos + numPuntosGuardar);
        }
    } catch (const std::bad_alloc &) {
        return THY_ERR_HOST_MEMORY;
    }
    double secs = 0.0;
    const int r = thyhost::run_problem(p, version ? version : "", &secs);
    if (tiempo) *tiempo = secs;
    return r;
}
