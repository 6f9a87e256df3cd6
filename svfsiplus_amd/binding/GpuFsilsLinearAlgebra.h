// GpuFsilsLinearAlgebra: the fluid Newton step of svFSIplus on one MI355X, as
// an implementation of the reference's linear-algebra seam.
//
// This file and GpuFsilsLinearAlgebra.cpp are written to sit beside the
// reference's own back-ends in Code/Source/solver (FsilsLinearAlgebra.h is
// the one they replace) and compile against its headers unchanged:
//
//   LinearAlgebra   (LinearAlgebra.h:40-63)       alloc, initialize, solve, ...
//   construct_fluid (fluid.h:51, fluid.cpp:464)   called at eq_assem.cpp:400-402
//   b_assem_neu_bc  (eq_assem.cpp:58-172)         called per Neumann face
//
// Everything below the seam is libsvgpu.so's C ABI (svfsi_gpu.h): the block
// pattern, R and Val live in device memory; per Newton iteration only Ag, Yg
// (and a face's hg) go up and the increment R comes down.
//
// Two ways to use it:
//  * solver only: the host assembles com_mod.R / com_mod.Val as before
//    (FsilsLinearAlgebra's contract, lhsa_ns::do_assem through assemble()),
//    solve() uploads them and runs fsils_solve on the device;
//  * device assembly: construct_fluid() and b_assem_neu_bc() of this class
//    replace the reference's for the fluid equation, and solve() uses the
//    device R / Val they filled (no host R / Val traffic).
// In the second mode, host code must not add to com_mod.R between the
// assembly and the solve (solve() checks that com_mod.R is still zero).
#ifndef GPU_FSILS_LINEAR_ALGEBRA_H
#define GPU_FSILS_LINEAR_ALGEBRA_H

#include "LinearAlgebra.h"
#include "svfsi_gpu.h"

#include <map>

class GpuFsilsLinearAlgebra : public LinearAlgebra {
  public:
    // `type` is the interface type the object reports: fsils (a drop-in for
    // FsilsLinearAlgebra) unless the caller has added an enumerator of its own.
    explicit GpuFsilsLinearAlgebra(consts::LinearAlgebraType type = consts::LinearAlgebraType::fsils);
    ~GpuFsilsLinearAlgebra() override;

    void alloc(ComMod& com_mod, eqType& lEq) override;
    void assemble(ComMod& com_mod, const int num_elem_nodes, const Vector<int>& eqN, const Array3<double>& lK,
                  const Array<double>& lR) override;
    void check_options(const consts::PreconditionerType prec_cond_type,
                       const consts::LinearAlgebraType assembly_type) override;
    void initialize(ComMod& com_mod, eqType& lEq) override;
    void set_assembly(consts::LinearAlgebraType assembly_type) override;
    void set_preconditioner(consts::PreconditionerType prec_type) override;
    void solve(ComMod& com_mod, eqType& lEq, const Vector<int>& incL, const Vector<double>& res) override;

    // Device replacements of fluid::construct_fluid and eq_assem::b_assem_neu_bc
    // for the fluid equation (TET4 P1-P1, constant viscosity, one domain,
    // fixed mesh; TRI3 faces).
    void construct_fluid(ComMod& com_mod, const mshType& lM, const Array<double>& Ag, const Array<double>& Yg);
    void b_assem_neu_bc(ComMod& com_mod, const faceType& lFa, const Vector<double>& hg, const Array<double>& Yg);

    // Options beyond the reference (svg_solve_opts: mixed precision, the
    // learned and two-level preconditioners); the default is the reference's
    // FP64 solve.
    void set_solve_options(const svg_solve_opts& opts) { opts_ = opts; }
    const svg_solve_opts& solve_options() const { return opts_; }

    // true after a Newton iteration assembled on the device
    bool device_assembly() const { return device_assembly_; }

  private:
    void ensure_device(ComMod& com_mod);
    void upload_state(ComMod& com_mod, const Array<double>& Ag, const Array<double>& Yg, bool withAg);
    void zero_if_pending(ComMod& com_mod);

    svg_lhs* lhs_ = nullptr;
    svg_msh* msh_ = nullptr;
    const mshType* mshOf_ = nullptr;
    std::map<const faceType*, svg_face*> faces_;
    double* dR_ = nullptr;
    double* dVal_ = nullptr;
    double* dAg_ = nullptr;
    double* dYg_ = nullptr;
    double* dHg_ = nullptr;
    double* dBf_ = nullptr;
    double* hStage_ = nullptr;  // pinned staging of Ag / Yg (4, tnNo)
    int nNo_ = 0, nnz_ = 0;
    bool zeroPending_ = true;       // the device R / Val are cleared before the next assembly
    bool device_assembly_ = false;  // this Newton iteration assembled on the device
    bool device_mode_ = false;      // some iteration assembled on the device
    svg_solve_opts opts_{};
};

#endif
