function [clstack,corrstack,shift_equations,shift_equations_map,clstack_mask]=...
    cryo_clmatrix_hip(pf,NK,verbose,max_shift,shift_step,map_filter_radius,...
    ref_clmatrix,ref_shifts_2d)
% CRYO_CLMATRIX_HIP  Common lines of all pairs of images on an MI355X GPU.
%
% Drop-in for cryo_clmatrix (sinograms/cryo_clmatrix.m): same arguments,
% defaults and outputs, computed by libclm through the clm_mex gateway.
%   NK                is accepted and ignored, as cryo_clmatrix_cpu.m:123 does.
%   verbose           bit 1: libclm's stage reports; bit 2 with ref_clmatrix:
%                     clstack_mask marks the pairs whose lines are within 5
%                     degrees of ref_clmatrix (cryo_clmatrix_cpu.m:567-617).
%   map_filter_radius > 0 filters every correlation map with
%                     cryo_average_clmap before the search (FP64).
%   ref_shifts_2d     only feeds the reference's per-pair debug printing; it is
%                     accepted and not used.
% The search runs in FP16 on the matrix cores, re-scored in FP32 and, for
% near ties, in FP64 (see clm_mex.c for the other precisions).

if ~exist('verbose','var') || isempty(verbose), verbose=1; end
if ~exist('max_shift','var') || isempty(max_shift), max_shift=15; end
if ~exist('shift_step','var') || isempty(shift_step), shift_step=1; end
if ~exist('map_filter_radius','var') || isempty(map_filter_radius), map_filter_radius=0; end
if ~exist('ref_clmatrix','var') || isscalar(ref_clmatrix), ref_clmatrix=[]; end
precision=2;   % CLM_FP16_RESCORE

[clstack,corrstack,shift_equations,shift_equations_map,clstack_mask]=...
    clm_mex(pf,verbose,max_shift,shift_step,precision,map_filter_radius,ref_clmatrix);
