function [b] = mobilityPrecon(f, swimmers, B, BT, A1, A2, tol, maxIter) %#ok<INUSD>
%MOBILITYPRECON Block preconditioner of the mobility problem on the GPU
%   (Mobility/MobilityPrecon.m:17-39): R^-1 f by the leaf-block solve
%   (reducedSolve), then out = -A2 \ (BT.' S4 (B.' (A1 \ UOm))) with S4 the
%   coronaRes-4 KIFMM, and R^-1 BT out subtracted.  tol and maxIter are
%   unused, as in the reference.
b = nnkifmm_mex('precon', nnkMobility(swimmers, B, BT, A1, A2, true), f);
end
