# Builds the reference library (RIKEN-RCCS/EigenG-Batched) from its own sources
# where they lie, for gfx950, into oracle/_ref/ (git-ignored, test use only).
# Nothing from the reference is copied into this repository: only the built
# shared object lands in oracle/_ref/.  The reference's own Makefiles are not
# used; this recipe compiles its two translation units directly:
#   eigen_GPU_batch.cu  (the solver, C ABI eigen_GPU_batch_FP/_DP)
#   eigen_GPU_check.cu  (its FP64 accuracy checker, eigen_GPU_check_FP/_DP)
# -D__GFX9__ selects the wave64 build (gpu_arch.h:44-49).
REF     ?= /root/reference
HIPCC   ?= /opt/rocm/bin/hipcc
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_ref
FLAGS   := -O3 --offload-arch=gfx950 -D__GFX9__ -fPIC -std=c++17 -w \
           -Xclang -target-feature -Xclang -packed-fp32-ops \
           -fhip-fp32-correctly-rounded-divide-sqrt

all: $(OUT)/libeigenGbatch_ref.so

$(OUT)/libeigenGbatch_ref.so: $(REF)/eigen_GPU_batch.cu $(REF)/eigen_GPU_check.cu
	mkdir -p $(OUT)
	$(HIPCC) $(FLAGS) -I$(REF) -x hip -c $(REF)/eigen_GPU_batch.cu -o $(OUT)/eigen_GPU_batch.o
	$(HIPCC) $(FLAGS) -I$(REF) -x hip -c $(REF)/eigen_GPU_check.cu -o $(OUT)/eigen_GPU_check.o
	$(HIPCC) -shared --offload-arch=gfx950 -o $@ $(OUT)/eigen_GPU_batch.o $(OUT)/eigen_GPU_check.o
	rm -f $(OUT)/eigen_GPU_batch.o $(OUT)/eigen_GPU_check.o

.PHONY: all
