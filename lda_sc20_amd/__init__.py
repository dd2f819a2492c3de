"""MI355X-native EZLDA Gibbs-sampling sweep.

The product is the HIP library libezlda.so (C ABI: include/ezlda.h), built
in-tree by lda_sc20_amd.build; lda_sc20_amd.ezlda is its ctypes binding.
"""
