function [out, flag, relres, iter, resvec] = nnkSolveMobility(swimmers, fluidFunc, t, conditions, ...
    kernalPar, TreePar, threads, GPU, blockSize, tol, maxIter, PreconSettings)
%NNKSOLVEMOBILITY The body of mobilityProblem(Gravity) after the bodies
%   are placed (Mobility/mobilityProblem.m:31-72): tree, KIFMM, B/BT,
%   right-hand side [u_body - u_fluid (format 2); F; M], GMRES on the GPU
%   (nnkifmm_mex('solve')) in place of MATLAB's gmres, and the unpacking
%   of U and Omega.
swimmers.genTree(TreePar{:});
swimmers.genKIFMM(kernalPar, 'parThreads', threads, 'GPU', GPU, 'GMRES', 1, 'format', 2, ...
                  'blockSize', blockSize);
[B, BT, A1, A2] = genMatrixComponets(swimmers);
pointVel = swimmers.getVelocities();
fluidVel = fluidFunc(t, swimmers.getSwimmers());
vel = pointVel;
vel(1:size(fluidVel, 1), :) = vel(1:size(fluidVel, 1), :) - fluidVel;
[force, moment] = swimmers.getForceMoment();
input = [vel(:); force(1:3:end); force(2:3:end); force(3:3:end); ...
         moment(1:3:end); moment(2:3:end); moment(3:3:end)];
precon = ~isempty(PreconSettings);
[x, flag, relres, iter, resvec] = nnkifmm_mex('solve', nnkMobility(swimmers, B, BT, A1, A2, precon), ...
                                              input, tol, maxIter, precon);
nsw = swimmers.swimmerNo();
x0 = x(end - 6 * nsw + 1:end - 3 * nsw);
omega = x(end - 3 * nsw + 1:end);
x0 = reshape(reshape(x0, [], 3).', [], 1);
omega = reshape(reshape(omega, [], 3).', [], 1);
out = zeros(size(conditions));
for i = 0:nsw - 1
    out((i * 9) + 1:(i * 9) + 3) = x0((i * 3) + 1:(i * 3) + 3);
    out((i * 9) + 4:(i * 9) + 6) = cross(omega((i * 3) + 1:(i * 3) + 3), swimmers.b{i + 1, 1});
    out((i * 9) + 7:(i * 9) + 9) = cross(omega((i * 3) + 1:(i * 3) + 3), swimmers.b{i + 1, 2});
end
end
