/*
 * ref_dos_harness.c -- drives the reference's own dos_interp
 * (dos_interpolate.c:107-124, DOS build), compiled from its source under
 * /root/reference into oracle/_ref/libttcref_dos.so, so that the oracle's
 * orc_dos_interp can be checked against it.
 *
 * TEST INFRASTRUCTURE ONLY (never shipped, never measured as the product).
 *
 * Fills box 0's Wang-Landau window (sim_dos[0]) and its bins (dos_hist[0])
 * the way init.c:340-346 does, then calls dos_interp.
 */
#define MAIN
#include "defines.h"

double dos_interp(int ibox, int bin, double val);

double ref_dos_interp(int nbins, const double *lng, double e_begin, double e_width, int bin, double val)
{
    if (nbins < 1 || nbins > NBINS || bin < 0 || bin >= nbins) return 0.0 / 0.0;
    sim_dos[0].e_begin = e_begin;
    sim_dos[0].e_width = e_width;
    sim_dos[0].e_bins = nbins;
    sim_dos[0].e_end = e_begin + e_width * nbins;
    for (int i = 0; i < nbins; i++) {
        dos_hist[0][i].e_mid = sim_dos[0].e_begin + sim_dos[0].e_width * (i + 0.5);   /* init.c:345 */
        dos_hist[0][i].g_of_e = lng[i];
        dos_hist[0][i].h_of_e = 0;
    }
    return dos_interp(0, bin, val);
}
