# Builds the reference's own solver sources (Niceno/PSI-BOIL, read where they
# lie under REF) into oracle/_ref/, plus this repo's driver ref_driver.cpp.
# Test infrastructure only: the output checks the C oracle and serves as the
# "reference" CPU baseline.  Nothing is copied out of REF.
#
# The reference's sequential configuration (first:74-80, "NOMPI") links
# Src/Parallel/mpi_macros.h -> sequential.h before building.  The source tree
# is read-only here, so the same symlink is made under oracle/_ref/cfg/ and
# the include search path resolves "../Parallel/mpi_macros.h" (and the
# deeper relative forms) to it.
#
#   make -f oracle/ref.mk -j8        (from the repo root)

REF     ?= /root/reference/Src
OUT     := oracle/_ref
CXX     ?= g++
CXXFLAGS:= -O2 -std=c++11 -ffp-contract=off -w -fpermissive

# Src/Makefile.am SUBDIRS (the configured build)
SUBDIRS := Equation/Staggered/Momentum Equation/Staggered Solver/Additive \
  Solver/Linear/Gauss Solver/Linear/Iterative Solver/Linear/Krylov Solver/Linear \
  Solver/Preconditioner Equation/Centered/ElectPoten Equation/Centered/Enthalpy \
  Equation/Centered/EnthalpyFD/EnthalpyFDaxisym Equation/Centered/EnthalpyFD \
  Equation/Centered/Concentration Equation/Centered/ConcentrationTP \
  Equation/Centered/CavityPressure/CPaxisym Equation/Centered/CavityPressure \
  Equation/Centered/Pressure Equation/Centered/PhaseChange4 Equation/Centered/CIPCSL2 \
  Equation/Centered/VOF/VOFaxisym Equation/Centered/VOF Equation/Centered/Distance \
  Equation/Centered Equation/CommonHeatTransfer Equation/Dispersed Equation/Floodfill \
  Equation/Heaviside/MarchingCubes Equation/Heaviside/MarchingSquares/MSaxisym \
  Equation/Heaviside/MarchingSquares Equation/Heaviside Equation/Nucleation/Microlayer \
  Equation/Nucleation Equation/Tifmodel/Antoine Equation/Tifmodel/Clapeyron \
  Equation/Tifmodel Equation/Topology Monitor/Rack Monitor/Location Matter LookUpTable \
  Grid Plot Plot/TEC Plot/VTK Custom/IF97 Custom Domain/Axisymmetric Domain Body Model \
  Field/Vector/TwoLevelVector Field/Vector Field/Scalar/TwoLevelScalar Field/Scalar \
  Field/ScalarBool Field/VectorBool Field/ScalarInt Timer Formula Profile RandomFlow \
  Global Boundary Ravioli SimulationTime Parallel/Out Parallel Template

# the sources each directory's Makefile.am lists (stale files beside them are
# not part of the reference's build)
SRCS := $(foreach d,$(SUBDIRS),$(addprefix $(REF)/$(d)/,$(shell grep -o '[A-Za-z0-9_.+-]*\.cpp' $(REF)/$(d)/Makefile.am)))
OBJS := $(patsubst $(REF)/%.cpp,$(OUT)/obj/%.o,$(SRCS))

CFG  := $(OUT)/cfg
INCS := -I$(CFG)/Parallel -I$(CFG)/a -I$(CFG)/a/b -I$(CFG)/a/b/c

all: $(OUT)/ref_driver $(OUT)/gk_driver

$(CFG)/.stamp:
	mkdir -p $(CFG)/Parallel $(CFG)/a/b/c
	ln -sf $(REF)/Parallel/sequential.h $(CFG)/Parallel/mpi_macros.h
	touch $@

$(OUT)/obj/%.o: $(REF)/%.cpp $(CFG)/.stamp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INCS) -c $< -o $@

$(OUT)/libboil.a: $(OBJS)
	rm -f $@
	ar rcs $@ $(OBJS)

$(OUT)/oracle_c.o: oracle/psb_oracle.c oracle/psb_oracle.h
	gcc -O2 -std=c11 -ffp-contract=off -c $< -o $@

$(OUT)/ref_driver: oracle/ref_driver.cpp $(OUT)/libboil.a $(OUT)/oracle_c.o
	$(CXX) $(CXXFLAGS) $(INCS) -I$(REF) -Ioracle $< $(OUT)/oracle_c.o \
	  -Wl,--start-group $(OUT)/libboil.a -Wl,--end-group -o $@ -lm

# the reference's callers of Linear::solve with the GpuKrylov binding
# (include/psb_krylov.hpp) over libpsb.so (built first by psi-boil_amd/build.py);
# the rpath finds the library from the binary's own directory
LIBPSB := psi-boil_amd/libpsb.so
$(OUT)/gk_driver: oracle/gk_driver.cpp oracle/ref_driver.cpp include/psb_krylov.hpp include/psb.h \
                  $(OUT)/libboil.a $(OUT)/oracle_c.o $(LIBPSB)
	$(CXX) $(CXXFLAGS) $(INCS) -I$(REF) -Ioracle -Iinclude $< $(OUT)/oracle_c.o \
	  -Wl,--start-group $(OUT)/libboil.a -Wl,--end-group -Lpsi-boil_amd -lpsb \
	  -Wl,-rpath,'$$ORIGIN/../../psi-boil_amd' -o $@ -lm

.PHONY: all
