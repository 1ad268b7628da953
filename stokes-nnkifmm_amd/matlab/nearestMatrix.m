function [NN] = nearestMatrix(coarse, fine, blockSize, threads) %#ok<INUSD>
%NEARESTMATRIX NN(q, j) = 1/#ties for the coarse point(s) j nearest to
%   fine point q (Tools/Nearest/nearestMatrix.m:17-41), by an exact grid
%   search on the GPU.  blockSize and threads are accepted and unused.
NN = nnkifmm_mex('nearest_matrix', coarse.', fine.');
end
