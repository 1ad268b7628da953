classdef OcTree < handle
    %OCTREE Adaptive octree and U/V/W/X interaction lists, built by
    %   libnnkifmm on the GPU with the semantics of the reference's
    %   Classes/OcTree.m (same ids, corners, split rule and list order).
    %
    %   tree = OcTree(potPoints, targetPoints, 'parThreads', p,
    %                 'nodeCapacity', c, 'maxDepth', d, 'finePoints', F,
    %                 'nearest', 0|1, 'NN', NN, 'blockSize', b)
    %   One option beyond the reference: 'rootCube', [lo hi] fixes the
    %   root box (time stepping with operator reuse).
    properties
        points; potentials; targets; finePoints; NN;
        pointIndex; nodeCorners; nodeLevel; nodeParents; nodeChildren;
        nodeCount; nodeCapacity; interactions; arguments;
    end
    properties (Hidden)
        h;   % nnkHandle of the library tree
    end
    methods
        function this = OcTree(potPoints, targetPoints, varargin)
            IP = inputParser;
            addParameter(IP, 'parThreads', 0);
            addParameter(IP, 'nodeCapacity', 200);
            addParameter(IP, 'maxDepth', 21);
            addParameter(IP, 'finePoints', []);
            addParameter(IP, 'nearest', 0);
            addParameter(IP, 'NN', []);
            addParameter(IP, 'blockSize', 0.2);
            addParameter(IP, 'rootCube', []);
            parse(IP, varargin{:});
            this.arguments = IP.Results;
            this.h = nnkHandle(nnkifmm_mex('tree', potPoints.', targetPoints.', ...
                                           this.arguments.nodeCapacity, this.arguments.maxDepth, ...
                                           this.arguments.rootCube));
            [this.points, pot, tgt, pidx, this.nodeCorners, lev, par, ch] = ...
                nnkifmm_mex('tree_export', this.h.id);
            this.potentials = logical(pot);
            this.targets = logical(tgt);
            this.pointIndex = pidx;
            this.nodeLevel = lev.';
            this.nodeParents = par.';
            this.nodeChildren = ch;
            this.nodeCount = numel(lev);
            this.nodeCapacity = min(this.arguments.nodeCapacity, size(this.points, 1) - 1);
            this.interactions = cell(this.nodeCount, 4);
            for kind = 0:3
                [ptr, idx] = nnkifmm_mex('tree_list', this.h.id, kind);
                for n = 1:this.nodeCount
                    this.interactions{n, kind + 1} = idx(ptr(n) + 1:ptr(n + 1)).';
                end
            end
            if this.arguments.nearest && ~isempty(this.arguments.finePoints)
                this.finePoints = this.arguments.finePoints;
                this.arguments.finePoints = [];
            end
            if isempty(this.arguments.NN) && this.arguments.nearest
                this.NN = nnkifmm_mex('nearest_matrix', potPoints.', this.finePoints.');
            else
                this.NN = this.arguments.NN;
                this.arguments.NN = [];
            end
            if this.arguments.nearest
                nnkifmm_mex('set_nearest', this.h.id, this.finePoints.', sparse(this.NN));
            end
        end
        function out = getPotentialPoints(this)
            out = this.points(this.potentials, :);
        end
        function out = getTargetPoints(this)
            out = this.points(this.targets, :);
        end
        function out = getFinePoints(this)
            out = this.finePoints;
        end
    end
end
