/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Harness around the reference's own parameter reader: links
 * read_param.c, confirm_param.c and msg.c, compiled from the reference tree
 * by oracle/ref.mk into oracle/_ref/ref_params.  Usage:
 *
 *     ref_params <confirm 0|1> <param file>
 *
 * prints the parsed Parameters as "key=value" lines after the line
 * "--- parameters", or ends through the reference's msg_abort (MPI_Abort with
 * the abort code) exactly as the reference driver would.
 */
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>

#include "parameters.h"
#include "msg.h"

static void put_string(const char* key, const char* s, int len)
{
  if (s == 0)
    printf("%s=<absent> %d\n", key, len);
  else
    printf("%s=[%s] %d\n", key, s, len);
}

int main(int argc, char* argv[])
{
  MPI_Init(&argc, &argv);
  msg_init();
  if (argc < 3) {
    fprintf(stderr, "usage: ref_params <confirm 0|1> <param file>\n");
    MPI_Finalize();
    return 2;
  }
  const int confirm = atoi(argv[1]);
  Parameters param;
  read_parameters(argc, argv, &param);
  if (confirm) {
    msg_set_loglevel(100); /* checks only, no listing */
    confirm_parameters(&param, argv[argc - 1]);
  }
  printf("--- parameters\n");
  printf("nc=%d\npm_nc_factor=%d\nnp_alloc_factor=%.17g\nntimestep=%d\n", param.nc, param.pm_nc_factor,
         param.np_alloc_factor, param.ntimestep);
  printf("random_seed=%d\nnrealization=%d\nloglevel=%d\n", param.random_seed, param.nrealization, param.loglevel);
  printf("a_final=%.17g\nboxsize=%.17g\nomega_m=%.17g\nsigma8=%.17g\nh=%.17g\n", param.a_final, param.boxsize,
         param.omega_m, param.sigma8, param.h);
  printf("n_zout=%d\n", param.n_zout);
  for (int i = 0; i < param.n_zout; i++) printf("zout=%.17g\n", param.zout[i]);
  put_string("power_spectrum_filename", param.power_spectrum_filename, param.strlen_power_spectrum_filename);
  put_string("fof_filename", param.fof_filename, param.strlen_fof_filename);
  put_string("snapshot_filename", param.snapshot_filename, param.strlen_snapshot_filename);
  put_string("subsample_filename", param.subsample_filename, param.strlen_subsample_filename);
  put_string("cgrid_filename", param.cgrid_filename, param.strlen_cgrid_filename);
  put_string("init_filename", param.init_filename, param.strlen_init_filename);
  printf("fof_linking_factor=%.17g\nsubsample_factor=%.17g\ncgrid_nc=%d\nwrite_longid=%d\n",
         param.fof_linking_factor, param.subsample_factor, param.cgrid_nc, param.write_longid);
  fflush(stdout);
  MPI_Finalize();
  return 0;
}
