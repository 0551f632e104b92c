# Builds the unmodified reference TeaLeaf (Fortran 90 + one C timer) straight
# from its sources under $(REF) into oracle/_ref/.  Test infrastructure only:
# the binary is the CPU checker and CPU baseline, never part of the product.
#
#   make -f oracle/ref.mk            # -> oracle/_ref/tea_leaf
#
# Toolchain: AMD flang (ROCm) + the MPICH headers/libraries that ship in
# /opt/conda of this image.  Nothing from the reference is copied: objects,
# module files and the executable are written to oracle/_ref/ only.

REF    ?= /root/reference
ORACLE := $(patsubst %/,%,$(dir $(lastword $(MAKEFILE_LIST))))
OUT    := $(ORACLE)/_ref
REF_FC ?= /opt/rocm/llvm/bin/amdflang
CC_REF ?= gcc
MPIDIR ?= /opt/conda
FFLAGS ?= -O3 -fopenmp
CFLAGS_REF ?= -O3

# Compile order follows the reference's module dependencies (makefile.deps).
SRCS := data.f90 definitions.f90 global_mpi.f90 report.f90 \
        kernels/pack_kernel.f90 pack.f90 tea.f90 timer.f90 parse.f90 \
        read_input.f90 kernels/initialise_chunk_kernel.f90 initialise_chunk.f90 \
        build_field.f90 kernels/update_halo_kernel.f90 \
        kernels/update_internal_halo_kernel.f90 update_halo.f90 \
        kernels/set_field_kernel.f90 set_field.f90 start.f90 \
        kernels/generate_chunk_kernel.f90 generate_chunk.f90 initialise.f90 \
        kernels/field_summary_kernel.f90 field_summary.f90 calc_dt.f90 \
        timestep.f90 kernels/tea_leaf_common_kernel.f90 tea_leaf_common.f90 \
        kernels/tea_leaf_cheby_kernel.f90 tea_leaf_cheby.f90 \
        kernels/tea_leaf_ppcg_kernel.f90 tea_leaf_ppcg.f90 \
        kernels/tea_leaf_cg_kernel.f90 tea_leaf_cg.f90 \
        kernels/tea_leaf_jacobi_kernel.f90 tea_leaf_jacobi.f90 tea_solve.f90 \
        visit.f90 diffuse.f90 tea_leaf.f90

OBJS := $(addprefix $(OUT)/,$(notdir $(SRCS:.f90=.o)))

all: $(OUT)/tea_leaf

$(OUT):
	mkdir -p $(OUT)

# Serial on purpose: each object needs the .mod files of the ones before it.
$(OUT)/tea_leaf: | $(OUT)
	@set -e; for s in $(SRCS); do \
	  o=$(OUT)/$$(basename $${s%.f90}).o; \
	  if [ ! -f $$o ] || [ $(REF)/$$s -nt $$o ]; then \
	    echo "FC $$s"; \
	    $(REF_FC) $(FFLAGS) -I$(MPIDIR)/include -module-dir $(OUT) -c $(REF)/$$s -o $$o 2>>$(OUT)/build.log; \
	  fi; \
	done
	$(CC_REF) $(CFLAGS_REF) -c $(REF)/timer_c.c -o $(OUT)/timer_c.o
	$(REF_FC) $(FFLAGS) $(OBJS) $(OUT)/timer_c.o -L$(MPIDIR)/lib -Wl,-rpath,$(MPIDIR)/lib \
	  -lmpifort -lmpi -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
