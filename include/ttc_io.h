/*
 * ttc_io.h -- the reference's on-disk inputs and trajectory output, read and
 * written natively so that reference-format systems drive the device path.
 *
 * Part of libttc.so (host C, no GPU calls).  Every reader keeps the
 * reference's token rules (whitespace-separated fields, 4-character names,
 * per-line record counts) and its parameter-matching rules, so the arrays it
 * returns equal what the reference's own readers leave in its globals for the
 * same files:
 *
 *   INPUT/simul.input   read_sim.c:36-440 (keyword form: sim_* and box_*)
 *   INPUT/mol.input     read_config.c:34-150
 *   INPUT/<param file>  read_atom.c:28-510 (atom types, bond/bend/torsion/
 *                       improper/Urey-Bradley parameters, LJ and 1-4 LJ,
 *                       NBFIX)
 *   INPUT/simul.psf     read_topology.c:26-891 (sites, bonds, bends,
 *                       torsions, impropers, donors, acceptors, explicit
 *                       exclusions; 1-4 pairs derived from bend pairs; the
 *                       STYPE exclusion table xint)
 *   INPUT/simul<k>.crd  read_config.c:236-325 (coordinates, name check,
 *                       pbc_init.c:21-66 wrapping)
 *   LJ tables           setlj.c:181-214 (STYPE, Lorentz-Berthelot, NBFIX)
 *   traj .trr frames    write_trr.c:22-200 (GROMACS trr, big-endian,
 *                       8-byte reals rounded through float, nm)
 *
 * Return codes are the reference's exit codes for the same condition: 1 a
 * file missing, a count or name mismatch or a parameter not found; 2 a
 * missing required simul.input value; 102 a psf site-count mismatch; 11 an
 * allocation failure; 111 a bad argument.  Messages: ttc_io_last_error().
 */
#ifndef TTC_IO_H
#define TTC_IO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ttc_io ttc_io;

/* Sizes of what was read (one box). */
typedef struct {
    int nsite;    /* box[k].boxns                                          */
    int ncomp;    /* sim.NC, molecule types in mol.input                   */
    int nbond;    /* bondN[k]                                              */
    int nbend;    /* bendN[k]                                              */
    int nurey;    /* ureyN[k] (Urey-Bradley pairs found for the bends)     */
    int ntors;    /* torsN[k]                                              */
    int nimpr;    /* imprN[k]                                              */
    int nexcl;    /* exNBN[k] (explicit NNB exclusions)                    */
    int n14;      /* in14N[k]                                              */
    int nxint;    /* xintN[k] = nbond + nbend + nexcl + n14 (S14 build)    */
    int ntype;    /* distinct atom types (atomids) present in the box      */
} ttc_io_counts;

/* The simulation settings the nonbonded path needs, from simul.input. */
typedef struct {
    double lx, ly, lz;          /* box_size_xyz                   */
    double rc, rclist;          /* sim_cutoff, sim_cutoff_nblist  */
    double rc_ew, rclist_ew;    /* sim_cutoff_ewald(_nblist)      */
    double kappa;               /* box_kappa                      */
    double eps_rf;              /* box_dielectric (1 if absent)   */
    double dt;                  /* sim_timestep [ps]              */
    double temp;                /* box_temp [K]                   */
    double nhc_q;               /* box_nh_tstat_mass              */
    int order;                  /* box_spme_spline_order          */
    int grid[3];                /* box_grid_pts_xyz               */
    int nhc_m;                  /* box_nh_tstat_n                 */
    int respa;                  /* box_respa                      */
    int nboxes;                 /* sim_nboxes                     */
    int mode, intmode;          /* sim_mode, sim_intmode          */
} ttc_io_sim;

/* Read box ibox of a reference input directory (the folder holding
 * simul.input, mol.input, the parameter file named by sim_paramfile,
 * simul.psf and simul<ibox>.crd; the reference reads them from ./INPUT). */
int ttc_io_read(const char *input_dir, int ibox, ttc_io **out);

void ttc_io_free(ttc_io *s);
const char *ttc_io_last_error(void);

int ttc_io_get_counts(const ttc_io *s, ttc_io_counts *c);
int ttc_io_get_sim(const ttc_io *s, ttc_io_sim *sim);

/* Per-site arrays (any pointer may be NULL).  pos_nopbc: coordinates as read
 * (atnopbc); pos: wrapped into [-L/2, L/2) as pbc_init does (atom); both
 * n x 3.  type: dense index into the ttc_io_get_types tables.  names: 5
 * bytes per site (NUL-terminated site name from the psf). */
int ttc_io_get_sites(const ttc_io *s, double *pos_nopbc, double *pos, double *q, double *mass,
                     int *atomid, int *type, int *resid, char *names);

/* LJ tables over the ntype types present, ordered by ascending atomid:
 * atomid[t]; eps, sig, eps14, sig14 are ntype x ntype row-major, the
 * reference's ljset[a][b].eps/.sig and ljset14 (setlj.c:181-205). */
int ttc_io_get_types(const ttc_io *s, int *atomid, double *eps, double *sig, double *eps14, double *sig14);

/* Bonded terms in the reference's units and order (read_topology.c):
 * bonds ab[2*i], req, k (kJ/mol/A^2, E = k (r - req)^2);
 * bends abc[3*i], angeq [rad], k; ureys ab[2*i], s0, k;
 * torsions abcd[4*i] with six Fourier slots each (kphi [kJ/mol], nphi,
 * delphi [rad]; unused slots zero); impropers abcd[4*i], k, angeq [rad]. */
int ttc_io_get_bonds(const ttc_io *s, int *ab, double *req, double *k);
int ttc_io_get_bends(const ttc_io *s, int *abc, double *angeq, double *k);
int ttc_io_get_ureys(const ttc_io *s, int *ab, double *s0, double *k);
int ttc_io_get_torsions(const ttc_io *s, int *abcd, double *kphi, double *nphi, double *delphi);
int ttc_io_get_impropers(const ttc_io *s, int *abcd, double *k, double *angeq);

/* Explicit exclusions (a, b) in NNB order; 1-4 quads (a, b, c, d) in the
 * order read_topology.c:739-809 finds them; the STYPE exclusion table xint
 * (a, b, xflag: 1 bond, 2 bend a-c, 3 explicit, 4 1-4 a-d), as
 * read_topology.c:847-891 fills it. */
int ttc_io_get_exclusions(const ttc_io *s, int *ab);
int ttc_io_get_14(const ttc_io *s, int *abcd);
int ttc_io_get_xint(const ttc_io *s, int *a, int *b, int *xflag);

/* Append one frame to a GROMACS .trr file exactly as write_trr.c writes it:
 * header (magic 1993, version 13, empty title, section sizes), box
 * (L/10 nm on the diagonal), then positions, velocities and forces, each
 * divided by 10, rounded to float and stored as big-endian doubles.  A NULL
 * x, v or f leaves that section out (write_trr.c writes x only).  Arrays are
 * (pointer to the first site's x, stride in doubles), so &atom[k][0].x with
 * stride 8 or &vv[k][0].x with stride 3 pass straight through.
 * time_ind and time_val are the frame's step index and time [ps], as
 * write_trr.c:72-84 computes them from the cycle. */
int ttc_io_write_trr(const char *path, int natoms, const double *x, int xstride, const double *v, int vstride,
                     const double *f, int fstride, double lx, double ly, double lz, long time_ind, double time_val);

#ifdef __cplusplus
}
#endif

#endif /* TTC_IO_H */
