/*
 * graphConfig.h -- run configuration passed to every algorithm entry point.
 *
 * Field order and types follow the reference's struct Arguments
 * (02_capi_graph/include/config/graphConfig.h:11-52) so that callers compiled against
 * the reference headers fill the same fields. The fields the PageRank/SpMV pull path
 * reads are: iterations, epsilon, pushpull, ker_numThreads and the afu/cu config words
 * (accepted and ignored: there is no FPGA to configure).
 */
#ifndef GRAPHCONFIG_H
#define GRAPHCONFIG_H

#include <stdint.h>
#include "mt19937.h"

#define WEIGHTED 0
#define DIRECTED 1

struct Arguments
{
    int wflag;
    int xflag;
    int sflag;
    int Sflag;
    int dflag;
    uint32_t binSize;
    uint32_t verbosity;
    uint32_t iterations;
    uint32_t trials;
    double epsilon;
    int source;
    uint32_t algorithm;
    uint32_t datastructure;
    uint32_t pushpull;
    uint32_t sort;
    uint32_t lmode;
    uint32_t lmode_l2;
    uint32_t lmode_l3;
    uint32_t mmode;
    uint32_t symmetric;
    uint32_t weighted;
    uint32_t delta;
    int pre_numThreads;
    int algo_numThreads;
    int ker_numThreads;
    char *fnameb;
    char *fnamel;
    uint32_t fnameb_format;
    uint32_t convert_format;
    mt19937state mt19937var;
    uint32_t cache_size;

    /* accelerator configuration words (FPGA AFU/CU bit-fields in the reference) */
    uint64_t afu_config;
    uint64_t afu_config_2;
    uint64_t cu_config;
    uint64_t cu_config_2;
    uint64_t cu_config_3;
    uint64_t cu_config_4;
};

#endif
