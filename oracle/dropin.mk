# The drop-in at the reference's plugin point: the reference's own driver
# (tea_solve.f90, diffuse.f90, timestep.f90) and set-up (tea.f90, start.f90,
# initialise.f90, read_input.f90, parse.f90, build_field.f90, ...) compiled
# unmodified from their sources under $(REF), linked with this repository's
# driver-wrapper modules (tealeaf_ref_amd/dropin: update_halo_module,
# set_field_module, tea_leaf_common/cg/cheby/ppcg/jacobi_module,
# generate_chunk, field_summary), whose kernel branches call
# libtealeaf_hip.so.  The reference's kernel files are not compiled (except
# initialise_chunk_kernel and pack_kernel, which its set-up and MPI layer
# call).  Nothing from the reference is copied: objects, module files and
# the executable go to oracle/_ref/dropin/ only.
#
#   make -f oracle/dropin.mk           # -> oracle/_ref/dropin/tea_leaf
#
# Needs libtealeaf_hip.so built first (make -C tealeaf_ref_amd/csrc).

REF    ?= /root/reference
ORACLE := $(patsubst %/,%,$(dir $(lastword $(MAKEFILE_LIST))))
ROOT   := $(ORACLE)/..
OUT    := $(ORACLE)/_ref/dropin
PKG    := $(ROOT)/tealeaf_ref_amd
REF_FC ?= /opt/rocm/llvm/bin/amdflang
CC_REF ?= gcc
MPIDIR ?= /opt/conda
FFLAGS ?= -O2

# In module-dependency order; ours: paths under tealeaf_ref_amd/.
SRCS := $(REF)/data.f90 $(REF)/definitions.f90 $(REF)/global_mpi.f90 $(REF)/report.f90 \
        $(REF)/kernels/pack_kernel.f90 $(REF)/pack.f90 $(REF)/tea.f90 $(REF)/timer.f90 \
        $(REF)/parse.f90 $(REF)/read_input.f90 $(REF)/kernels/initialise_chunk_kernel.f90 \
        $(REF)/initialise_chunk.f90 $(REF)/build_field.f90 \
        $(PKG)/host/tl_hip_api.f90 $(PKG)/host/tl_eigen.f90 $(PKG)/dropin/tl_dropin.f90 \
        $(PKG)/dropin/tl_dropin_solvers.f90 $(PKG)/dropin/tl_dropin_setup.f90 \
        $(REF)/start.f90 $(REF)/initialise.f90 $(REF)/calc_dt.f90 $(REF)/timestep.f90 \
        $(REF)/tea_solve.f90 $(REF)/visit.f90 $(REF)/diffuse.f90 $(REF)/tea_leaf.f90

all: $(OUT)/tea_leaf

$(OUT):
	mkdir -p $(OUT)

# Serial on purpose: each object needs the .mod files of the ones before it.
$(OUT)/tea_leaf: $(SRCS) $(PKG)/libtealeaf_hip.so | $(OUT)
	@set -e; objs=""; for s in $(SRCS); do \
	  o=$(OUT)/$$(basename $${s%.f90}).o; objs="$$objs $$o"; \
	  echo "FC $$s"; \
	  $(REF_FC) $(FFLAGS) -I$(MPIDIR)/include -module-dir $(OUT) -c $$s -o $$o 2>>$(OUT)/build.log; \
	done; \
	$(CC_REF) -O2 -c $(REF)/timer_c.c -o $(OUT)/timer_c.o; \
	$(REF_FC) $(FFLAGS) $$objs $(OUT)/timer_c.o -L$(PKG) -ltealeaf_hip \
	  -Wl,-rpath,'$$ORIGIN/../../../tealeaf_ref_amd' -L$(MPIDIR)/lib -Wl,-rpath,$(MPIDIR)/lib \
	  -lmpifort -lmpi -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
