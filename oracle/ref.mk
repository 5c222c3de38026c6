# Builds the reference CPU corrector (care-cpu) from its own sources where
# they lie under /root/reference, into oracle/_ref/.  This is test
# infrastructure: the binary is the parity oracle and the CPU baseline, never
# part of the product path.
#
# The reference's include/helpers is a symlink into an un-vendored git
# submodule (hpc_helpers).  Its vendored copy inside warpcore is used
# instead, reached through a symlink made inside oracle/_ref/inc/ (nothing is
# written to /root/reference and no header is written by hand).
#
# Flags follow the reference Makefile (Makefile:21-24): -O3 -march=native
# -fopenmp, thrust's OpenMP backend.
#
#   make -f oracle/ref.mk            (from the repository root)

REF      ?= /root/reference
OUT      ?= oracle/_ref
CXX      ?= g++
JOBS     ?= 8

SRCS = correct_cpu.cpp correctionresultoutput.cpp cpu_alignment.cpp \
       cpuminhasherconstruction.cpp dispatch_care_correct_cpu.cpp \
       main_correct_cpu.cpp msa.cpp options.cpp readlibraryio.cpp threadpool.cpp

OBJS = $(SRCS:%.cpp=$(OUT)/obj/%.o)

CXXFLAGS_REF = -std=c++17 -Wno-terminate -Wno-class-memaccess -w -fopenmp \
               -I$(OUT)/inc -I$(REF)/include -I$(REF)/dependencies/thrust-1.17.0 \
               -O3 -march=native -DTHRUST_DEVICE_SYSTEM=THRUST_DEVICE_SYSTEM_OMP

.PHONY: all clean
all: $(OUT)/care-cpu

$(OUT)/inc/helpers:
	mkdir -p $(OUT)/inc $(OUT)/obj
	ln -sfn $(REF)/dependencies/warpcore/ext/hpc_helpers/include $(OUT)/inc/helpers

$(OUT)/obj/%.o: $(REF)/src/%.cpp | $(OUT)/inc/helpers
	$(CXX) $(CXXFLAGS_REF) -c $< -o $@

$(OUT)/care-cpu: $(OBJS)
	$(CXX) $(OBJS) -lpthread -lgomp -lstdc++fs -lz -ldl -o $@

clean:
	rm -rf $(OUT)/obj $(OUT)/inc $(OUT)/care-cpu
