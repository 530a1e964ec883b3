# Build recipe for oracle/_ref: the reference's own FP64 Fortran core, compiled
# from the sources where they lie (REFSRC, read-only) with amdflang, plus this
# project's driver (oracle/ref_driver.f90) that exposes the reference routines
# through bind(C) entry points for ctypes.  Test infrastructure only: nothing
# under specfab_amd/ links or loads it.
#
# LAPACK: the reference links the system -llapack (src/Makefile:7); this image
# has none, so the reference objects' LAPACK references (dsyev_, dposv_,
# dgetrf_, dgetri_) are renamed with objcopy to the LAPACK entry points that
# scipy's bundled OpenBLAS 0.3.28 exports (scipy_dsyev_ ...), and the library
# links against that .so.  No reference source is copied or modified.
#
#   make -f oracle/ref.mk            (from the repo root)

REFSRC  ?= /root/reference/src
OUT     ?= oracle/_ref
FC      := /opt/rocm/bin/amdflang
FFLAGS  ?= -O2 -fPIC -m64 -w
OBJ      = $(OUT)/obj
OPENBLAS_DIR ?= $(shell python3 -c "import scipy,os;print(os.path.join(os.path.dirname(scipy.__file__),'..','scipy.libs'))")
OPENBLAS_SO  ?= $(firstword $(wildcard $(OPENBLAS_DIR)/libscipy_openblas*.so))

MODS = header tensorproducts mandel reducedform gaunt lambdasolver rheologies \
       elasticities moments frames rotation idealstate dynamics homogenizations \
       enhancementfactors wavepropagation deformationmodes
M77  = amach derv1 dnqsol dnrm2 erfin ermsg ierm1 ierv1

MODOBJS = $(addprefix $(OBJ)/,$(addsuffix .o,$(MODS)))
M77OBJS = $(addprefix $(OBJ)/m77_,$(addsuffix .o,$(M77)))
RENAMES = --redefine-sym dsyev_=scipy_dsyev_ --redefine-sym dposv_=scipy_dposv_ \
          --redefine-sym dgetrf_=scipy_dgetrf_ --redefine-sym dgetri_=scipy_dgetri_ \
          --redefine-sym dgeev_=scipy_dgeev_

all: $(OUT)/libsfref.so

$(OBJ):
	mkdir -p $(OBJ)

# Modules are compiled in dependency order (each .o depends on the previous one
# so that the .mod files exist); -J puts .mod files beside the objects.
$(OBJ)/header.o: $(REFSRC)/header.f90 | $(OBJ)
	$(FC) $(FFLAGS) -J$(OBJ) -I$(REFSRC) -c $< -o $@
	objcopy $(RENAMES) $@

define MODRULE
$(OBJ)/$(2).o: $(REFSRC)/$(2).f90 $(OBJ)/$(1).o
	$(FC) $(FFLAGS) -J$(OBJ) -I$(REFSRC) -c $(REFSRC)/$(2).f90 -o $(OBJ)/$(2).o
	objcopy $(RENAMES) $(OBJ)/$(2).o
endef
$(eval $(call MODRULE,header,tensorproducts))
$(eval $(call MODRULE,tensorproducts,mandel))
$(eval $(call MODRULE,mandel,reducedform))
$(eval $(call MODRULE,reducedform,gaunt))
$(eval $(call MODRULE,gaunt,lambdasolver))
$(eval $(call MODRULE,lambdasolver,rheologies))
$(eval $(call MODRULE,rheologies,elasticities))
$(eval $(call MODRULE,elasticities,moments))
$(eval $(call MODRULE,moments,frames))
$(eval $(call MODRULE,frames,rotation))
$(eval $(call MODRULE,rotation,idealstate))
$(eval $(call MODRULE,idealstate,dynamics))
$(eval $(call MODRULE,dynamics,homogenizations))
$(eval $(call MODRULE,homogenizations,enhancementfactors))
$(eval $(call MODRULE,enhancementfactors,wavepropagation))
$(eval $(call MODRULE,wavepropagation,deformationmodes))

$(OBJ)/m77_%.o: $(REFSRC)/include/math77/%.f | $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/ref_driver.o: oracle/ref_driver.f90 $(OBJ)/deformationmodes.o
	$(FC) $(FFLAGS) -J$(OBJ) -I$(OBJ) -c $< -o $@

$(OUT)/libsfref.so: $(OBJ)/ref_driver.o $(MODOBJS) $(M77OBJS)
	$(FC) -shared -fPIC -o $@ $^ -Wl,--no-undefined $(OPENBLAS_SO) -Wl,-rpath,$(OPENBLAS_DIR) -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
