function [est_shifts,shift_equations]=cryo_estimate_shifts_hip(pf,rotations,...
    max_shift,shift_step,memoryfactor,shifts_2d_ref,verbose)
% CRYO_ESTIMATE_SHIFTS_HIP  2-D shifts of the images from their rotations on
% an MI355X GPU.
%
% Drop-in for cryo_estimate_shifts (sinograms/cryo_estimate_shifts.m): same
% arguments, defaults and outputs, computed by libclm through the
% clm_estimate_shifts_mex gateway.  The common line of each pair from the
% rotations, the 1-D shift along it and the least squares (lsqr, tol 1e-8,
% maxit 100) run on the GPU in FP64; shift_equations is assembled on the
% host in FP64.
%   memoryfactor   picks the number of equations as the reference does; the
%                  pairs are a uniform draw seeded with 0 (the reference's
%                  randperm stream is not reproduced) and the equations come
%                  in pair order.
%   shifts_2d_ref, verbose  only feed the reference's diagnostics; they are
%                  accepted and not used.

if ~exist('shift_step','var') || isempty(shift_step), shift_step=1; end
if ~exist('memoryfactor','var') || isempty(memoryfactor), memoryfactor=10000; end
seed=0;

[est_shifts,shift_equations]=...
    clm_estimate_shifts_mex(pf,rotations,max_shift,shift_step,memoryfactor,seed);
