classdef KIFMM
    %KIFMM Kernel-independent FMM for regularized Stokeslets on the GPU
    %   (libnnkifmm), with the options and vector layouts of the
    %   reference's Classes/KIFMM.m.
    %
    %   fmm = KIFMM(tree, [eps mu], 'GPU', g, 'parThreads', p, 'blockSize', b,
    %               'GMRES', 0|1, 'format', 1|2, 'coronaRes', r,
    %               'coronaShells', s)
    %   Options beyond the reference: 'precision' ('fp64' (default),
    %   'mixed', 'mixed3', 'mixedh', 'fp64t') and 'svdTol' (pseudo-inverse
    %   cut of the truncated forms).  GPU, parThreads and blockSize are accepted and
    %   unused: the library always runs on the GPU.
    properties
        tree;
        arguments;
    end
    properties (Hidden)
        ops;   % nnkOperators shared by copies of this object
    end
    methods
        function this = KIFMM(tree, kernelPar, varargin)
            this.tree = tree;
            IP = inputParser;
            addParameter(IP, 'GPU', 0);
            addParameter(IP, 'parThreads', 0);
            addParameter(IP, 'blockSize', 0.1);
            addParameter(IP, 'GMRES', 0);
            addParameter(IP, 'coronaRes', 6);
            addParameter(IP, 'coronaShells', 1);
            addParameter(IP, 'format', 1);
            addParameter(IP, 'precision', 'fp64');
            addParameter(IP, 'svdTol', 0);
            parse(IP, varargin{:});
            this.arguments = IP.Results;
            this.arguments.kernelPar = kernelPar;
            this.ops = nnkOperators();
            this.handle(this.arguments.coronaRes);   % build the operator now, as the reference sets up here
        end
        function vel = computeVel(this, potentials)
            vel = this.run('apply', this.arguments.coronaRes, potentials);
        end
        function vel = computeVelSparse(this, potentials)
            vel = this.run('apply', 4, potentials);   % coronaRes 4 (KIFMM.m:136-145)
        end
        function vel = reducedComputeVel(this, potentials)
            vel = this.run('reduced', this.arguments.coronaRes, potentials);
        end
        function pot = reducedSolve(this, vel)
            pot = this.run('rsolve', this.arguments.coronaRes, vel);
        end
        function id = handle(this, res)
            p = find(strcmp(this.arguments.precision, {'fp64', 'mixed', 'mixed3', 'mixedh', 'fp64t'})) - 1;
            if isempty(p)
                error('KIFMM: unknown precision ''%s''', this.arguments.precision);
            end
            shells = this.arguments.coronaShells;
            if res == 4
                shells = 1;
            end
            id = this.ops.get(this.tree, this.arguments.kernelPar, res, shells, p, this.arguments.svdTol);
        end
    end
    methods (Access = private)
        function out = run(this, op, res, in)
            % layouts of KIFMM.m:83-103: (n,3), or with GMRES = 1 a column
            % [x1 y1 z1 ...]' (format 1) or [x1..xN y1..yN z1..zN]' (format 2)
            if this.arguments.GMRES == 1
                if this.arguments.format == 1
                    in = reshape(in, 3, []).';
                elseif this.arguments.format == 2
                    in = reshape(in, [], 3);
                end
            end
            out = nnkifmm_mex(op, this.handle(res), in.').';
            if this.arguments.GMRES == 1
                if this.arguments.format == 1
                    out = reshape(out.', [], 1);
                elseif this.arguments.format == 2
                    out = reshape(out, [], 1);
                end
            end
        end
    end
end
