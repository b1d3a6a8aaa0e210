/* The rsf command line as a C entry point (librsf_cli.so).
 *
 * rsf_cli_main(argc, argv) is the whole rsf executable -- the drop-in for the
 * reference's rsf3d (CUDA_RSF/main.cpp:73-319): same long options, .npy inputs
 * and output, per-iteration stdout line and exit codes.  The executable's main
 * only calls it.  A process that already uses the GPU can run the command line
 * through it without starting a child process. */
#ifndef RSF_CLI_H
#define RSF_CLI_H

#ifdef __cplusplus
extern "C" {
#endif

int rsf_cli_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif

#endif
