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
#   make -f oracle/ref.mk -j8

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
              eq_assem heats heatf l_elas cep
SOLVER_SRC := $(addprefix $(SOLVER)/,$(addsuffix .cpp,$(SOLVER_TUS)))

LS_OBJ     := $(patsubst $(LS)/%.cpp,$(OBJ)/ls_%.o,$(LS_SRC))
SOLVER_OBJ := $(patsubst $(SOLVER)/%.cpp,$(OBJ)/sv_%.o,$(SOLVER_SRC))
HARN_OBJ   := $(OBJ)/ref_harness.o

all: $(OUT)/libsvref.so

$(OBJ):
	mkdir -p $(OBJ)

$(OBJ)/ls_%.o: $(LS)/%.cpp | $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJ)/sv_%.o: $(SOLVER)/%.cpp | $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(HARN_OBJ): $(HERE)ref_harness.cpp | $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/exports.map: | $(OBJ)
	printf '{ global: ref_*; local: *; };\n' > $@

$(OUT)/libsvref.so: $(LS_OBJ) $(SOLVER_OBJ) $(HARN_OBJ) $(OUT)/exports.map
	$(CXX) -shared -o $@ $(HARN_OBJ) $(LS_OBJ) $(SOLVER_OBJ) \
	  -Wl,--gc-sections -Wl,--version-script=$(OUT)/exports.map -Wl,--no-undefined \
	  -static-libstdc++ -static-libgcc \
	  $(CONDA)/lib/libmpi.so $(CONDA)/lib/libmkl_rt.so -Wl,-rpath,$(CONDA)/lib

.PHONY: all
