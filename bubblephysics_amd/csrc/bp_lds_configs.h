// bp_lds_configs.h -- the launch configurations of the LDS-resident world kernel.
#ifndef BP_LDS_CONFIGS_H
#define BP_LDS_CONFIGS_H

// Launch configurations of the LDS kernel: threads, bubbles per thread, LDS row capacity, and
// the segment, body, vertex and triangle capacities of the staged geometry.
#define BP_LDS_CONFIGS(X)                          \
    X(0, 64, 1, 1024, 32, 32, 128, 128)            \
    X(1, 128, 1, 2048, 32, 32, 128, 128)           \
    X(2, 256, 1, 4096, 64, 64, 256, 256)           \
    X(3, 512, 1, 8192, 64, 128, 512, 512)          \
    X(4, 1024, 1, 16384, 64, 128, 512, 512)        \
    X(5, 256, 2, 8192, 64, 64, 256, 256)           \
    X(6, 512, 2, 16384, 64, 128, 512, 512)         \
    X(7, 1024, 2, 24576, 64, 128, 512, 512)        \
    X(8, 256, 4, 16384, 64, 64, 256, 256)          \
    X(9, 1024, 4, 0, 32, 32, 128, 128)

#endif  // BP_LDS_CONFIGS_H
