/*
 * semsks_uvw2sac -- command-line drop-in for `uvw2sac par=run.par`
 * (semsyn/plot/uvw2sac.c): seis.u/.v/.w of a finished SEM run to SAC files.
 * Exit statuses follow the reference's getpar (100 missing parameter, 101
 * STOP) and open_file (255) conventions.
 */
#include <stdio.h>
#include "semsks.h"

int main(int argc, char **argv)
{
    semsks_status st = semsks_uvw2sac(argc, argv);
    if (st == SEMSKS_OK) return 0;
    fprintf(stderr, "semsks_uvw2sac: %s\n", semsks_last_error());
    switch (st) {
    case SEMSKS_EPAR: return 100;
    case SEMSKS_ESTOP: return 101;
    case SEMSKS_EIO: return 255;
    default: return 1;
    }
}
