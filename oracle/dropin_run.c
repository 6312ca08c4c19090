/*
 * dropin_run.c -- an NVE run of the reference's own program pieces on the
 * GPU path: reference-format inputs, the reference's integrator, its .trr
 * writer and its A/B restart files.
 *
 * TEST INFRASTRUCTURE ONLY.  Linked (by `make -C oracle dropin`) from the
 * reference's unmodified read_sim.c, read_atom.c, read_config.c,
 * read_topology.c, allocate.c, setlj.c, pbc_init.c, pbc_all.c, forces.c,
 * cbond.c, cbend.c, curey.c, ctorsion.c, cimproper.c, integrate_vverlet.c,
 * kinet.c, write_trr.c and checkpoint.c, compiled from /root/reference with
 * the flags NLIST DLIST STYPE S14 EWALD PRESSURE TRR CHECKP, together with
 * thecodecombo_7_amd/host/ttc_refshim.c and libttc.so.
 *
 * Usage: dropin_run CASE_DIR VEL.bin OUT.bin nsteps mode ckpt_step restart
 *                   trr_every chunk
 *   CASE_DIR  holds INPUT/ (read by the reference's readers in main.c's
 *             order: read_sim, read_atom, read_config, setlj), and receives
 *             CHECKP/ and OUTPUT/GROMACS/
 *   VEL.bin   float64 vel[3n] (the start velocities, A/ps)
 *   mode 0    nve_md.c:360-390: nl_check / nblist, integrate_vverlet(k, 1)
 *             and kinet(k) every step, each force call through the shim
 *   mode 1    the same steps in device-resident chunks of `chunk` steps
 *             (ttc_shim_md_vverlet), the host state brought up to date
 *             after each chunk
 *   ckpt_step after this step: WriteCheckpoint(equiname[checkptr], ...)
 *             then checkptr = 1 - checkptr (nve_md.c:152-160), and a list
 *             rebuild (so a restarted run and this one share the anchors)
 *   restart 1 InitCheckpoint finds the restart files; ReadCheckpoint(
 *             equiname[checkptr], ...) then nblist(k), and the run continues
 *             from the step recorded in the file with the restored forces
 *   trr_every write_trr(k, step, 0) every trr_every steps (0: never)
 *   OUT.bin   float64 atom[3n], atnopbc[3n], vv[3n], en.kinet, en.potens
 *             (calcvalue.c:42-53 sum), then int32 last step
 */
#define MAIN
#include "defines.h"

#include <sys/stat.h>
#include <unistd.h>

void read_sim(void);
void read_atom(void);
void read_config(void);
void setlj(void);
void ewald_setup(int);
void nblist(int);
void nl_check(int, int, int);
void forces(int);
void integrate_vverlet(int, double);
void kinet(int);
void write_trr(int, int, int);
void pme_clean(void);
void InitCheckpoint(void);
int WriteCheckpoint(const char *filename, unsigned long i, double vscal[NBOXES], int state);
int ReadCheckpoint(const char *filename, int *i, double vscal[NBOXES], int state);
void ttc_shim_md_vverlet(int k, int nsteps, double vscal);
extern int FLAG_x, FLAG_v, FLAG_f;

int main(int argc, char **argv)
{
    if (argc != 10) {
        fprintf(stderr, "usage: dropin_run CASE_DIR VEL.bin OUT.bin nsteps mode ckpt_step restart trr_every chunk\n");
        return 2;
    }
    const char *outp = argv[3];
    int nsteps = atoi(argv[4]), mode = atoi(argv[5]), ckpt = atoi(argv[6]), restart = atoi(argv[7]);
    int trr_every = atoi(argv[8]), chunk = atoi(argv[9]);
    char velp[4096], outabs[4096];
    if (!realpath(argv[2], velp)) { perror(argv[2]); return 2; }
    snprintf(outabs, sizeof outabs, "%s", outp);
    if (outp[0] != '/') {
        char cwd[4096];
        if (!getcwd(cwd, sizeof cwd)) return 2;
        snprintf(outabs, sizeof outabs, "%s/%s", cwd, outp);
    }
    if (chdir(argv[1])) { perror(argv[1]); return 2; }
    int k = 0;
    read_sim();
    read_atom();
    read_config();
    setlj();
    int n = box[k].boxns;
    FILE *f = fopen(velp, "rb");
    if (!f) { perror(velp); return 2; }
    for (int i = 0; i < n; i++)
        if (fread(&vv[k][i].x, sizeof(double), 3, f) != 3) { fprintf(stderr, "short VEL.bin\n"); return 2; }
    fclose(f);
    for (int i = 0; i < n; i++) uu[k][i] = vv[k][i];
    fileindex = 0;
    FLAG_x = 1; FLAG_v = 1; FLAG_f = 1;
    mkdir("OUTPUT", 0755);
    mkdir("OUTPUT/GROMACS", 0755);
    InitCheckpoint();
    double vscal[NBOXES] = {1.0};
    int start = 0;
    ewald_setup(k);
    if (restart) {
        /* checkptr names the newest file, which is also the next one written
         * (checkpoint.c:135-137); the restored ff[k] is the force the next
         * step's first half kick uses, as in the uninterrupted run */
        if (!restart_equil) { fprintf(stderr, "no restart files found\n"); return 3; }
        ReadCheckpoint(equiname[checkptr], &start, vscal, 0);
        nblist(k);
    } else {
        nblist(k);
        forces(k);
    }
    int step = start;
    while (step < nsteps) {
        int todo = 1;
        if (mode == 1) {
            todo = chunk;
            if (ckpt > step && step + todo > ckpt) todo = ckpt - step;
            if (trr_every) {
                int next = (step / trr_every + 1) * trr_every;
                if (step + todo > next) todo = next - step;
            }
            if (step + todo > nsteps) todo = nsteps - step;
            ttc_shim_md_vverlet(k, todo, 1.0);
        } else {
            nl_check(0, box[k].boxns, k);
            if (nl_flag[k] == 1) nblist(k);
            integrate_vverlet(k, 1.0);
        }
        step += todo;
        kinet(k);
        if (trr_every && step % trr_every == 0) write_trr(k, step, 0);
        if (step == ckpt) {
            WriteCheckpoint(equiname[checkptr], (unsigned long)step, vscal, 0);
            checkptr = 1 - checkptr;
            nblist(k);
        }
    }
    FILE *o = fopen(outabs, "wb");
    if (!o) { perror(outabs); return 2; }
    for (int i = 0; i < n; i++) fwrite(&atom[k][i].x, sizeof(double), 3, o);
    for (int i = 0; i < n; i++) fwrite(&atnopbc[k][i].x, sizeof(double), 3, o);
    for (int i = 0; i < n; i++) fwrite(&vv[k][i].x, sizeof(double), 3, o);
    double pot = en[k].bond + en[k].bend + en[k].tors + en[k].nbonds + en[k].impr + en[k].urey;
    fwrite(&en[k].kinet, sizeof(double), 1, o);
    fwrite(&pot, sizeof(double), 1, o);
    fwrite(&step, sizeof(int), 1, o);
    fclose(o);
    pme_clean();
    return 0;
}
