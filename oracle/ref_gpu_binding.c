/*
 * The reference-side binding of the MI355X engine, as a maintainer of the
 * reference adds it -- here only compiled and linked against the reference's
 * own sources (`make -C oracle ref-gpu`) to show that it is complete.
 *
 * It replaces src/correlator/src/gemm_gpu.c (the OpenCL path) with three
 * functions over the engine's C ABI, and gives the reference's CPU task
 * driver a per-task dispatch, so the `-gpu` run goes through the uint64
 * tables and the task loop of correlate_function.c (correlate_function.c:
 * 52-80, 239-290) instead of correlate_function_gpu.c, whose uint32 tables
 * misalign rows (correlate_IOx32.c:118-126 against the stride at
 * correlate_function_gpu.c:121-124) and whose task range drops every
 * thread's last task (correlate_function_gpu.c:55-59,76: with one thread and
 * one task it computes nothing).  The reference's edits, spelled out in the
 * recipe by compiler definitions so that no reference file is changed:
 *   pthreads.c:196,252       correlate_function_gpu(...) -> correlate_function(...)
 *   correlate_function.c     correlate(...) -> qld_task_correlate(...), same arguments
 *   gemm_gpu.c, correlate_function_gpu.c   not built
 * gpu_init / gpu_release keep their call sites (correlate_our_kernel.c:
 * 593-596, 626-629).  TEST INFRASTRUCTURE here: nothing in the engine uses
 * this file.
 */
#include <stdio.h>
#include <stdlib.h>

#include "header.h"
#include "gemm_cpu.h"
#include "qld_host.h"

static qld_ctx *g_qld;

void gpu_init(void)
{
    qld_cfg cfg;
    qld_cfg_default(&cfg);                 /* device 0, AUTO engine, REFERENCE r^2 */
    if (qld_gpu_init(&g_qld, &cfg) != QLD_OK) {
        fprintf(stderr, "qLD-compute: GPU init failed: %s\n", qld_gpu_last_error(NULL));
        exit(1);                           /* the command decides, not the library */
    }
}

void gpu_release(void)
{
    qld_gpu_release(g_qld);
    g_qld = NULL;
}

/* correlate_gpu (gemm_gpu.h:56-70) with the uint64 tables of the CPU task
 * driver: counts, r^2, selection and report text on the engine. */
static void correlate_gpu_x64(threadData_t *threadData, outFileType fpOut,
                              inputDataType_x64 *tableA, unsigned int *tableA_posIndex,
                              char **tableA_IDindex, unsigned int *tableA_bitcount,
                              int tableAsize, inputDataType_x64 *tableB,
                              unsigned int *tableB_posIndex, char **tableB_IDindex,
                              unsigned int *tableB_bitcount, int tableBsize,
                              int compressed_snp_size, int snp_size, int posWset2)
{
    const int rc = qld_correlate_report(g_qld, fpOut, tableA, tableA_posIndex, tableA_IDindex,
                                        tableA_bitcount, tableAsize, tableB, tableB_posIndex,
                                        tableB_IDindex, tableB_bitcount, tableBsize,
                                        compressed_snp_size, snp_size, posWset2,
                                        threadData[0].r2limit);
    if (rc != QLD_OK) {
        fprintf(stderr, "qLD-compute: GPU correlate failed (%d): %s\n", rc,
                qld_gpu_last_error(g_qld));
        exit(1);
    }
}

/* The task driver's per-task call: the GPU thread's tasks (the last thread,
 * pthreads.c:154-158) go to the engine, the other threads' to the CPU path
 * (gemm_cpu.c:420-623). */
void qld_task_correlate(threadData_t *threadData, outFileType fpOut,
                        inputDataType_x64 *tableA, unsigned int *tableA_posIndex,
                        char **tableA_IDindex, unsigned int *tableA_bitcount, int tableAsize,
                        inputDataType_x64 *tableB, unsigned int *tableB_posIndex,
                        char **tableB_IDindex, unsigned int *tableB_bitcount, int tableBsize,
                        int compressed_snp_size, int snp_size, int posWset2)
{
    if (threadData[0].gpu)
        correlate_gpu_x64(threadData, fpOut, tableA, tableA_posIndex, tableA_IDindex,
                          tableA_bitcount, tableAsize, tableB, tableB_posIndex, tableB_IDindex,
                          tableB_bitcount, tableBsize, compressed_snp_size, snp_size, posWset2);
    else
        correlate(threadData, fpOut, tableA, tableA_posIndex, tableA_IDindex, tableA_bitcount,
                  tableAsize, tableB, tableB_posIndex, tableB_IDindex, tableB_bitcount,
                  tableBsize, compressed_snp_size, snp_size, posWset2);
}
