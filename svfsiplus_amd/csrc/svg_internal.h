// Internal declarations shared by the host setup code and the HIP sources.
#pragma once

#include "../../include/svfsi_gpu.h"

#include <string>
#include <vector>

namespace svg {

int set_error(int code, const std::string& msg);

int pipe_counts(const svg_pipe_spec& sp, int* nNo, int* nEl);
int pipe_generate(const svg_pipe_spec& sp, double* x, int* IEN, int* flags);
int bif_counts(const svg_bif_spec& b, int* nNo, int* nEl);
int bif_generate(const svg_bif_spec& b, double* x, int* IEN, int* flags);
int csr_pattern(int nNo, int nEl, int eNoN, const int* IEN, std::vector<int>& rowPtr, std::vector<int>& colPtr);
int color_elements(int nNo, int nEl, int eNoN, const int* IEN, std::vector<int>& color, int* nColors);
int slot_map(int nEl, int eNoN, const int* IEN, const int* rowPtr, const int* colPtr, std::vector<int>& slots);

}  // namespace svg
