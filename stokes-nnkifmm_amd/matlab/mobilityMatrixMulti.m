function [b] = mobilityMatrixMulti(f, swimmers, B, BT)
%MOBILITYMATRIXMULTI [S f + BT [U; Omega]; B f] on the GPU
%   (Mobility/MobilityMatrixMulti.m:17-36): f = [force (3N, format 2);
%   U (3 nsw); Omega (3 nsw)].
b = nnkifmm_mex('mobility_apply', nnkMobility(swimmers, B, BT, [], [], false), f);
end
