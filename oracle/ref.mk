# Compiles the reference's own hot-path sources, where they lie under
# /root/reference, into oracle/_ref/libsvref.so together with
# oracle/ref_harness.cpp. Nothing from the reference is copied into the repo;
# the objects and the library stay under oracle/_ref (git-ignored).
#
# Requirements (all present in this image, nothing is stubbed):
#   g++ (C++17), MPICH headers/library and MKL (dgesv_) from /opt/conda.
# The reference's own build system (CMake + VTK) is not used. Symbols of the
# linked translation units that the harness never reaches (VTK readers pulled
# in by Simulation-based helpers) are removed by section garbage collection
# under a version script that exports only the harness's ref_* functions.
#
# The second target, oracle/_ref/libsvbind.so, compiles the drop-in back-end
# svfsiplus_amd/binding/GpuFsilsLinearAlgebra.cpp against the reference's own
# headers (LinearAlgebra.h, ComMod.h, ...) and links it, with
# oracle/bind_harness.cpp and the reference's solver/ls.cpp (ls_alloc,
# ls_solve), against the objects above and svfsiplus_amd/libsvgpu.so (built
# first by svfsiplus_amd/build.py); HIP's host runtime API from $(ROCM).
#
#   make -f oracle/ref.mk -j8            # both
#   make -f oracle/ref.mk -j8 bind

REF     ?= /root/reference
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_ref
OBJ     := $(OUT)/obj
SOLVER  := $(REF)/Code/Source/solver
LS      := $(REF)/Code/Source/liner_solver
CONDA   ?= /opt/conda

CXX      = g++
CXXFLAGS = -std=c++17 -O2 -fPIC -ffunction-sections -fdata-sections -w \
           -include cstring \
           -I$(SOLVER) -I$(LS) -I$(REF)/Code/ThirdParty/tinyxml/simvascular_tinyxml \
           -I$(REF)/Code/ThirdParty/eigen/include/eigen3 -I$(CONDA)/include

LS_SRC     := $(wildcard $(LS)/*.cpp)
SOLVER_TUS := fluid nn fs lhsa utils Array Array3 Vector all_fun ComMod CmMod consts CepMod \
              mat_fun Tensor4 FsilsLinearAlgebra LinearAlgebra PetscLinearAlgebra TrilinosLinearAlgebra \
              eq_assem heats heatf l_elas cep output Simulation Parameters CepModAp CepModBo CepModFn \
              CepModTtp ChnlMod
SOLVER_SRC := $(addprefix $(SOLVER)/,$(addsuffix .cpp,$(SOLVER_TUS)))

LS_OBJ     := $(patsubst $(LS)/%.cpp,$(OBJ)/ls_%.o,$(LS_SRC))
SOLVER_OBJ := $(patsubst $(SOLVER)/%.cpp,$(OBJ)/sv_%.o,$(SOLVER_SRC)) $(OBJ)/tinyxml2.o
TINYXML    := $(REF)/Code/ThirdParty/tinyxml/simvascular_tinyxml/tinyxml2.cpp
HARN_OBJ   := $(OBJ)/ref_harness.o

REPO    := $(abspath $(HERE)..)
BIND    := $(REPO)/svfsiplus_amd/binding
SVGLIB  := $(REPO)/svfsiplus_amd/libsvgpu.so
ROCM    ?= /opt/rocm
BIND_FLAGS = $(CXXFLAGS) -I$(BIND) -I$(REPO)/include -I$(ROCM)/include -D__HIP_PLATFORM_AMD__
BIND_OBJ := $(OBJ)/bind_GpuFsilsLinearAlgebra.o $(OBJ)/bind_harness.o $(OBJ)/sv_ls.o

all: $(OUT)/libsvref.so $(OUT)/libsvbind.so

bind: $(OUT)/libsvbind.so

$(OBJ):
	mkdir -p $(OBJ)

$(OBJ)/ls_%.o: $(LS)/%.cpp | $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJ)/sv_%.o: $(SOLVER)/%.cpp | $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJ)/tinyxml2.o: $(TINYXML) | $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(HARN_OBJ): $(HERE)ref_harness.cpp | $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJ)/bind_GpuFsilsLinearAlgebra.o: $(BIND)/GpuFsilsLinearAlgebra.cpp $(BIND)/GpuFsilsLinearAlgebra.h \
    $(REPO)/include/svfsi_gpu.h | $(OBJ)
	$(CXX) $(BIND_FLAGS) -c $< -o $@

$(OBJ)/bind_harness.o: $(HERE)bind_harness.cpp $(BIND)/GpuFsilsLinearAlgebra.h | $(OBJ)
	$(CXX) $(BIND_FLAGS) -c $< -o $@

$(OUT)/bind_exports.map: | $(OBJ)
	printf '{ global: bind_*; local: *; };\n' > $@

$(OUT)/libsvbind.so: $(BIND_OBJ) $(LS_OBJ) $(SOLVER_OBJ) $(SVGLIB) $(OUT)/bind_exports.map
	$(CXX) -shared -o $@ $(BIND_OBJ) $(LS_OBJ) $(SOLVER_OBJ) \
	  -Wl,--gc-sections -Wl,--version-script=$(OUT)/bind_exports.map -Wl,--no-undefined \
	  -static-libstdc++ -static-libgcc \
	  -L$(REPO)/svfsiplus_amd -lsvgpu -L$(ROCM)/lib -lamdhip64 \
	  $(CONDA)/lib/libmpi.so $(CONDA)/lib/libmkl_rt.so \
	  -Wl,-rpath,'$$ORIGIN/../../svfsiplus_amd' -Wl,-rpath,$(ROCM)/lib -Wl,-rpath,$(CONDA)/lib

$(OUT)/exports.map: | $(OBJ)
	printf '{ global: ref_*; local: *; };\n' > $@

$(OUT)/libsvref.so: $(LS_OBJ) $(SOLVER_OBJ) $(HARN_OBJ) $(OUT)/exports.map
	$(CXX) -shared -o $@ $(HARN_OBJ) $(LS_OBJ) $(SOLVER_OBJ) \
	  -Wl,--gc-sections -Wl,--version-script=$(OUT)/exports.map -Wl,--no-undefined \
	  -static-libstdc++ -static-libgcc \
	  $(CONDA)/lib/libmpi.so $(CONDA)/lib/libmkl_rt.so -Wl,-rpath,$(CONDA)/lib

.PHONY: all bind
